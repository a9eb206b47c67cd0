// Host runtime of the HIP MLS-MPM backend: the C ABI of include/tessera_mpm.h.
//
// Mirrors the reference's resident session (crates/tessera-mpm/src/gpu.rs:
// 206-662): the session owns the last synchronized f64 world (insertion
// order, the authoritative state), a device-resident f32 copy of the
// particles, and the per-substep topology.  submit_steps() enqueues up to 64
// substeps on the session's HIP stream without host round trips;
// synchronize() validates the device flags, converts the device state back
// to the f64 world and, on any failure, invalidates the session while keeping
// the last synchronized world (gpu.rs:377-411).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/tessera_mpm.h"
#include "mpm_device.h"
#include "mpm_kernels.h"

using namespace tmpm;

namespace {

thread_local std::string g_last_error;

tmpm_status fail(tmpm_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(TMPM_READBACK, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// Host-side restatement of the reference's validation rules.
// ---------------------------------------------------------------------------
inline bool finite3(const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// material.rs:429-431
bool valid_elasticity(double young, double poisson) {
  return std::isfinite(young) && young > 0.0 && std::isfinite(poisson) && poisson >= -1.0 && poisson < 0.5;
}
// material.rs:172-231
bool material_valid(const tmpm_material& m) {
  const double* p = m.p;
  switch (m.kind) {
    case TMPM_MATERIAL_LINEAR_ELASTIC:
    case TMPM_MATERIAL_NEO_HOOKEAN: return valid_elasticity(p[0], p[1]);
    case TMPM_MATERIAL_SAND:
    case TMPM_MATERIAL_SAND_NEO_HOOKEAN:
      return valid_elasticity(p[0], p[1]) && std::isfinite(p[2]) && p[2] >= 0.0 && p[2] < M_PI / 2 &&
             std::isfinite(p[3]) && p[3] >= 0.0;
    case TMPM_MATERIAL_FLUID:
      return std::isfinite(p[0]) && p[0] > 0 && std::isfinite(p[1]) && p[1] > 0 && std::isfinite(p[2]) && p[2] >= 0 &&
             std::isfinite(p[3]) && p[3] >= 0;
    case TMPM_MATERIAL_SNOW:
      return valid_elasticity(p[0], p[1]) && std::isfinite(p[2]) && p[2] >= 0 && p[2] < 1.0 && std::isfinite(p[3]) &&
             p[3] >= 0 && std::isfinite(p[4]) && p[4] >= 0;
  }
  return false;
}
// material.rs:433-438
void lame(double young, double poisson, double& lambda, double& mu) {
  lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  mu = young / (2.0 * (1.0 + poisson));
}
double det_cm(const double* F) {  // nalgebra 3x3 determinant on column-major storage
  double m11 = F[0], m21 = F[1], m31 = F[2], m12 = F[3], m22 = F[4], m32 = F[5], m13 = F[6], m23 = F[7], m33 = F[8];
  double a = m22 * m33 - m32 * m23, b = m21 * m33 - m31 * m23, c = m21 * m32 - m31 * m22;
  return m11 * a - m12 * b + m13 * c;
}
// material.rs:380-426
inline bool is_plastic(const tmpm_material& m) {
  return m.kind == TMPM_MATERIAL_SNOW || m.kind == TMPM_MATERIAL_SAND || m.kind == TMPM_MATERIAL_SAND_NEO_HOOKEAN;
}

double timestep_bound(const tmpm_material& m, double density, const double* v, double h, double det) {
  const double* p = m.p;
  double speed = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double lambda, mu;
  if (m.kind == TMPM_MATERIAL_SAND_NEO_HOOKEAN) {
    lame(p[0], p[1], lambda, mu);
    double d = density / std::max(det, 1e-6);
    double ws = std::sqrt((lambda + 2.0 * mu) / d);
    return 0.5 * h / std::max(ws, speed);
  }
  double wave;
  if (m.kind == TMPM_MATERIAL_FLUID) {
    wave = std::sqrt(p[0] * p[1] / std::max(density, 1e-12));
  } else {
    lame(p[0], p[1], lambda, mu);
    wave = std::sqrt((lambda + 2.0 * mu) / std::max(density, 1e-12));
  }
  return 0.4 * h / std::max(wave + speed, 1e-12);
}
// obstacle.rs:265-341
bool obstacle_valid(const tmpm_obstacle& o) {
  const double* d = o.dims;
  bool sv = false;
  switch (o.shape) {
    case TMPM_SHAPE_SPHERE: sv = std::isfinite(d[0]) && d[0] > 0; break;
    case TMPM_SHAPE_BOX:
      sv = std::isfinite(d[0]) && d[0] > 0 && std::isfinite(d[1]) && d[1] > 0 && std::isfinite(d[2]) && d[2] > 0;
      break;
    case TMPM_SHAPE_CAPSULE:
      sv = std::isfinite(d[0]) && d[0] >= 0 && std::isfinite(d[1]) && d[1] >= 0 && (d[0] > 0 || d[1] > 0);
      break;
    case TMPM_SHAPE_CYLINDER:
    case TMPM_SHAPE_CONE:
    case TMPM_SHAPE_GROUND: sv = std::isfinite(d[0]) && d[0] > 0 && std::isfinite(d[1]) && d[1] > 0; break;
    case TMPM_SHAPE_TRIANGLE_PRISM: {
      const double* v = o.vertices;
      double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
      double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      double a2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
      sv = std::isfinite(d[0]) && d[0] > 0 && finite3(v) && finite3(v + 3) && finite3(v + 6) && std::isfinite(a2) &&
           a2 > 1e-24;
      break;
    }
    case TMPM_SHAPE_CONVEX: {
      sv = o.plane_count >= 4 && o.planes != nullptr && std::isfinite(d[0]) && d[0] > 0;
      for (uint32_t i = 0; sv && i < o.plane_count; ++i) {
        const double* pl = o.planes + 4 * i;
        for (int k = 0; k < 4; ++k) sv = sv && std::isfinite(pl[k]);
        double n2 = std::fma(pl[0], pl[0], std::fma(pl[1], pl[1], pl[2] * pl[2]));
        sv = sv && std::fabs(n2 - 1.0) <= 1e-4;
      }
      break;
    }
    default: return false;
  }
  bool cpic_ok = o.cpic_group < 0 || (o.cpic_group < 32 && o.shape == TMPM_SHAPE_TRIANGLE_PRISM);
  return sv && cpic_ok && finite3(o.center) && std::isfinite(o.orientation[0]) && std::isfinite(o.orientation[1]) &&
         std::isfinite(o.orientation[2]) && std::isfinite(o.orientation[3]) && finite3(o.linear_velocity) &&
         finite3(o.angular_velocity) && std::isfinite(o.friction) && o.friction >= 0;
}
// world.rs:53-64
bool params_valid(const tmpm_params& p) {
  if (!std::isfinite(p.cell_width) || p.cell_width <= 0.0 || !finite3(p.gravity) || !std::isfinite(p.max_substep) ||
      p.max_substep <= 0.0)
    return false;
  if (p.has_bounds) {
    if (!finite3(p.bound_min) || !finite3(p.bound_max)) return false;
    for (int a = 0; a < 3; ++a)
      if (!(p.bound_max[a] - p.bound_min[a] > 0.0)) return false;
  }
  return true;
}

inline bool finite_f32(double v, float& out) {
  out = (float)v;
  return std::isfinite(out);
}

// f32 cell key with no fused multiply-add (gpu.rs:1142-1154).
inline bool cell_f32(float x, float inv_h, int32_t& c) {
  volatile float gp = x * inv_h;
  volatile float sh = gp - 0.5f;
  float f = std::floor((float)sh);
  if (!std::isfinite(f) || (double)f < -2147483648.0 || (double)f > 2147483645.0) return false;
  c = (int32_t)f;
  return true;
}

// ---------------------------------------------------------------------------
// Host world (f64, insertion order) — `MpmWorld` (world.rs:206-221).
// ---------------------------------------------------------------------------
// Runs fn(begin, end) over [0, n) in up to 16 host threads (chunks of at
// least min_chunk items; small ranges run inline).  Used for the per-particle
// host work of uploads and admissions.
template <class F>
void parallel_for(size_t n, F&& fn, size_t min_chunk = 16384) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t chunks = std::min<size_t>(hw, (n + min_chunk - 1) / min_chunk);
  if (chunks <= 1) {
    fn((size_t)0, n);
    return;
  }
  const size_t step = (n + chunks - 1) / chunks;
  std::vector<std::thread> pool;
  pool.reserve(chunks);
  size_t inline_from = chunks;  // chunks no thread could be started for run here
  for (size_t c = 1; c < chunks; ++c) {
    const size_t b = c * step, e = std::min(n, b + step);
    if (b >= e) break;
    try {
      pool.emplace_back([&fn, b, e] { fn(b, e); });
    } catch (const std::system_error&) {
      inline_from = c;
      break;
    }
  }
  fn((size_t)0, std::min(n, step));
  for (size_t c = inline_from; c < chunks; ++c)
    if (c * step < n) fn(c * step, std::min(n, c * step + step));
  for (auto& t : pool) t.join();
}

struct OwnedObstacle {
  tmpm_obstacle o;
  std::vector<double> planes;
};

// Allocator whose value-less construct default-initializes: resizing a
// vector of doubles leaves the new elements unwritten, so a parallel fill
// is the first touch of their pages.
template <class T>
struct DefaultInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() = default;
  template <class U>
  DefaultInit(const DefaultInit<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
using DVec = std::vector<double, DefaultInit<double>>;

struct HostWorld {
  tmpm_params params{};
  DVec pos, vel, F, C, force, mass, rest_volume, radius, damping, plastic;
  std::vector<tmpm_material> material;
  std::vector<uint32_t> group_id;
  std::vector<uint16_t> color;
  std::vector<uint8_t> enabled, fixed;
  std::vector<uint64_t> chunk;
  std::vector<OwnedObstacle> obstacles;
  uint64_t substeps = 0;
  uint64_t next_chunk_id = 1;

  size_t size() const { return mass.size(); }

  // Empties the world but keeps every vector's memory, for reuse by a later
  // admission (no fresh pages to fault in).
  void clear_keep() {
    for (auto* v : {&pos, &vel, &F, &C, &force, &mass, &rest_volume, &radius, &damping, &plastic}) v->clear();
    material.clear();
    group_id.clear();
    color.clear();
    enabled.clear();
    fixed.clear();
    chunk.clear();
    obstacles.clear();
    params = tmpm_params{};
    substeps = 0;
    next_chunk_id = 1;
  }
  size_t capacity_bytes() const { return pos.capacity() * 34 * sizeof(double) / 3; }

  void set_obstacles(const tmpm_obstacle* obs, uint32_t n) {
    obstacles.clear();
    obstacles.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      obstacles[i].o = obs[i];
      if (obs[i].shape == TMPM_SHAPE_CONVEX && obs[i].planes)
        obstacles[i].planes.assign(obs[i].planes, obs[i].planes + 4 * (size_t)obs[i].plane_count);
      obstacles[i].o.planes = obstacles[i].planes.empty() ? nullptr : obstacles[i].planes.data();
    }
  }
  HostWorld() = default;
  // moves keep every vector's buffer, so obstacle plane pointers stay valid
  HostWorld(HostWorld&&) noexcept = default;
  HostWorld& operator=(HostWorld&&) noexcept = default;
  HostWorld(const HostWorld& w) { *this = w; }
  HostWorld& operator=(const HostWorld& w) {
    if (this == &w) return *this;
    params = w.params;
    pos = w.pos; vel = w.vel; F = w.F; C = w.C; force = w.force; mass = w.mass; rest_volume = w.rest_volume;
    radius = w.radius; damping = w.damping; plastic = w.plastic; material = w.material; group_id = w.group_id;
    color = w.color; enabled = w.enabled; fixed = w.fixed; chunk = w.chunk;
    substeps = w.substeps; next_chunk_id = w.next_chunk_id;
    std::vector<tmpm_obstacle> raw;
    for (auto& o : w.obstacles) raw.push_back(o.o);
    set_obstacles(raw.data(), (uint32_t)raw.size());
    return *this;
  }

  // world.rs:492-515 (valid_particle, finite_particle_state) for particle i.
  bool particle_valid(size_t i) const {
    auto fin = [](const double* p, int n) {
      for (int k = 0; k < n; ++k)
        if (!std::isfinite(p[k])) return false;
      return true;
    };
    bool finite = fin(&pos[3 * i], 3) && fin(&vel[3 * i], 3) && fin(&F[9 * i], 9) && fin(&C[9 * i], 9) &&
                  fin(&force[3 * i], 3) && std::isfinite(mass[i]) && std::isfinite(rest_volume[i]) &&
                  std::isfinite(radius[i]) && std::isfinite(damping[i]);
    const double* pl = &plastic[3 * i];
    return finite && mass[i] > 0.0 && rest_volume[i] > 0.0 && radius[i] > 0.0 && material_valid(material[i]) &&
           damping[i] >= 0.0 && std::isfinite(pl[0]) && pl[0] > 0.0 && std::isfinite(pl[1]) && std::isfinite(pl[2]);
  }

  // world.rs:289-303
  double stable_timestep() const {
    const size_t n = size();
    std::vector<double> part((n + 16383) / 16384 + 1, params.max_substep);
    parallel_for(n, [&](size_t b, size_t e) {
      double s = params.max_substep;
      for (size_t i = b; i < e; ++i) {
        if (!enabled[i] || fixed[i]) continue;
        double t = timestep_bound(material[i], mass[i] / rest_volume[i], &vel[3 * i], params.cell_width,
                                  det_cm(&F[9 * i]));
        s = std::fmin(s, t);
      }
      part[b / 16384] = std::fmin(part[b / 16384], s);
    });
    double s = params.max_substep;
    for (double v : part) s = std::fmin(s, v);
    return std::fmin(s, params.max_substep);
  }

  // Appends a particle batch (world.rs:244-265); validation by the caller.
  void append(const tmpm_particles& p, const tmpm_material* mats, uint64_t chunk_id) {
    size_t n = p.count;
    // grow without writing, then copy (or fill) in parallel: the page faults
    // of fresh memory are taken by all threads
    auto app = [n](DVec& dst, const double* src, int k, const double* pattern) {
      const size_t base = dst.size(), m = n * k;
      dst.resize(base + m);
      double* d = dst.data() + base;
      parallel_for(m, [&](size_t b, size_t e) {
        if (src)
          std::memcpy(d + b, src + b, (e - b) * sizeof(double));
        else
          for (size_t i = b; i < e; ++i) d[i] = pattern[i % k];
      }, 1 << 16);
    };
    static const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static const double eye3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    static const double plastic0[3] = {1, 0, 0};
    app(pos, p.position, 3, zero9);
    app(vel, p.velocity, 3, zero9);
    app(F, p.deformation, 9, eye3);
    app(C, p.affine, 9, zero9);
    app(force, p.force, 3, zero9);
    app(mass, p.mass, 1, zero9);
    app(rest_volume, p.rest_volume, 1, zero9);
    app(radius, p.radius, 1, zero9);
    app(damping, p.damping, 1, zero9);
    app(plastic, p.plastic, 3, plastic0);
    const size_t base = material.size();
    material.resize(base + n);
    group_id.resize(base + n);
    color.resize(base + n);
    enabled.resize(base + n);
    fixed.resize(base + n);
    chunk.resize(base + n, chunk_id);
    parallel_for(n, [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        material[base + i] = mats[p.material_index ? p.material_index[i] : 0];
        group_id[base + i] = p.group_id ? p.group_id[i] : 0;
        color[base + i] = p.transfer_color ? p.transfer_color[i] : 0;
        enabled[base + i] = p.enabled ? (p.enabled[i] != 0) : 1;
        fixed[base + i] = p.fixed ? (p.fixed[i] != 0) : 0;
      }
    });
  }

  // world.rs:268-277: keep particles whose chunk differs.
  size_t remove_chunk(uint64_t id) {
    std::vector<size_t> keep;
    for (size_t i = 0; i < size(); ++i)
      if (chunk[i] != id) keep.push_back(i);
    size_t removed = size() - keep.size();
    if (removed == 0) return 0;
    auto sel = [&keep](auto& v, int k) {
      using V = typename std::remove_reference<decltype(v)>::type;
      V out;
      out.reserve(keep.size() * k);
      for (size_t i : keep)
        for (int q = 0; q < k; ++q) out.push_back(v[i * k + q]);
      v.swap(out);
    };
    sel(pos, 3); sel(vel, 3); sel(F, 9); sel(C, 9); sel(force, 3); sel(mass, 1); sel(rest_volume, 1);
    sel(radius, 1); sel(damping, 1); sel(plastic, 3); sel(material, 1); sel(group_id, 1); sel(color, 1);
    sel(enabled, 1); sel(fixed, 1); sel(chunk, 1);
    return removed;
  }
};

bool validate_particles(const tmpm_particles& p, const tmpm_material* mats, uint32_t mat_count) {
  if (p.count > 0 && (!p.position || !p.mass || !p.rest_volume || !p.radius)) return false;
  for (uint64_t i = 0; i < p.count; ++i)
    if (p.material_index && p.material_index[i] >= mat_count) return false;
  if (p.count > 0 && mat_count == 0) return false;
  HostWorld tmp;
  tmp.append(p, mats, 0);
  for (size_t i = 0; i < tmp.size(); ++i)
    if (!tmp.particle_valid(i)) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Device memory helpers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    release();
    bytes = b ? b : 16;
    return hipMalloc(&ptr, bytes);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct ParticleBuffers {
  DevBuf pos_mass, vel_vol, C, F, P, cls, ins_id;
};

struct KernelTimer {
  std::string name;
  hipEvent_t start, stop;
};

// Events are created once and reused: creating them per launch would add
// host latency to the measured stream.

}  // namespace

struct tmpm_context {
  int device = 0;
  hipDeviceProp_t props{};
};

// One independent world of a session.  A plain session has one scene; a
// batch steps several together, one launch per stage per substep, with the
// scene index in every block key's region so blocks never mix scenes.
struct SceneState {
  HostWorld world;           // last synchronized state, insertion order
  double substep_dt = 0;
  double max_mass = 0;       // largest enabled particle mass (fixed-point scales)
  bool force_pending = false;
  bool active = true;        // false: removed from the batch (slot free)
  tmpm_status status = TMPM_OK;  // error that stopped this scene in a batch
  DevScene dev{};
  uint32_t first = 0;        // insertion index of its particle 0 on the device
  uint32_t count = 0;        // particles on the device
  double stable = -1.0;      // world.stable_timestep(), cached; < 0 when stale
  bool device_newer = false; // batch: the device holds substeps not yet read into `world`
  // rigid coupling (bodies and one mapping per obstacle) and reactions
  std::vector<tmpm_rigid_body> bodies;
  std::vector<tmpm_rigid_mapping> maps;
  double rigid_gravity[3] = {0.0, 0.0, 0.0};
  double rigid_max_speed = 0.0;
  bool rigid_integrate = false;
  std::vector<double> reactions;  // 6 per obstacle, substeps of the last synchronize

  bool stepping() const { return active && status == TMPM_OK; }
  double stable_timestep() {
    if (stable < 0.0) stable = world.stable_timestep();
    return stable;
  }
  void set_world(const HostWorld& w) {
    world = w;
    stable = -1.0;
  }
  void set_world(HostWorld&& w) {
    world = std::move(w);
    stable = -1.0;
  }
};

// Growable host array of a trivially copyable T: resize leaves new elements
// uninitialized and clear keeps the memory, so a staging buffer reused
// across admissions touches no fresh pages once it has grown.  PINNED
// buffers are page-locked (hipHostMalloc) for full-rate uploads.
template <class T, bool PINNED = false>
struct PodBuf {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  PodBuf() = default;
  PodBuf(const PodBuf&) = delete;
  PodBuf& operator=(const PodBuf&) = delete;
  ~PodBuf() { release(p); }
  static T* acquire(size_t c) {
    void* q = nullptr;
    if (PINNED) {
      if (hipHostMalloc(&q, c * sizeof(T), hipHostMallocDefault) != hipSuccess) q = nullptr;
    } else {
      q = std::malloc(c * sizeof(T));
    }
    if (!q) throw std::bad_alloc();
    return static_cast<T*>(q);
  }
  static void release(T* q) {
    if (!q) return;
    if (PINNED) (void)hipHostFree(q);
    else std::free(q);
  }
  size_t size() const { return n; }
  bool empty() const { return n == 0; }
  T* data() { return p; }
  const T* data() const { return p; }
  T* begin() { return p; }
  T* end() { return p + n; }
  const T* begin() const { return p; }
  const T* end() const { return p + n; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
  void clear() { n = 0; }
  void reset() {
    release(p);
    p = nullptr;
    n = cap = 0;
  }
  void resize(size_t m) {
    if (m > cap) {
      const size_t c = std::max(m, cap + cap / 2);
      T* q = acquire(c);
      if (n) std::memcpy(q, p, n * sizeof(T));
      release(p);
      p = q;
      cap = c;
    }
    n = m;
  }
  void resize(size_t m, const T& v) {
    const size_t o = n;
    resize(m);
    for (size_t i = o; i < m; ++i) p[i] = v;
  }
};

// Host image of some scenes' enabled particles, ready for upload.
struct PackedParticles {
  PodBuf<float4, true> pm, vv;
  PodBuf<float> C, F;  // 9 per particle, particle-major
  PodBuf<float> Pl;    // plastic state, 3 per particle
  PodBuf<uint32_t, true> cls, ins;
  PodBuf<float> pos32, inv_h;  // for counting blocks
  PodBuf<uint32_t> region;
  PodBuf<float4> disabled;     // snapshot records of disabled particles
  PodBuf<float, true> cols;    // upload staging: component-major C and F
  PodBuf<float4, true> forces; // upload staging: one-shot forces
  size_t size() const { return pm.size(); }
  void clear() {
    pm.clear(); vv.clear(); C.clear(); F.clear(); Pl.clear(); cls.clear(); ins.clear();
    pos32.clear(); inv_h.clear(); region.clear(); disabled.clear(); cols.clear(); forces.clear();
  }
  void reset() {
    pm.reset(); vv.reset(); C.reset(); F.reset(); Pl.reset(); cls.reset(); ins.reset();
    pos32.reset(); inv_h.reset(); region.reset(); disabled.reset(); cols.reset(); forces.reset();
  }
};

// Interning key of a particle class: every DevClass field that varies.
using ClassKey = std::tuple<int, float, float, float, float, float, float, float, float, uint32_t, uint32_t, float,
                            float, float, float, float, float>;
inline ClassKey class_key(const DevClass& c) {
  return std::make_tuple(c.kind, c.radius, c.damping, c.lambda, c.mu, c.bulk, c.gamma, c.viscosity, c.tensile,
                         c.region, c.fixed, c.hardening, c.crit_compression, c.crit_stretch, c.friction_angle,
                         c.cohesion, c.inv_h);
}

// A captured substep: the launch plan's bytes and the instantiated graph.
struct CachedGraph {
  std::vector<unsigned char> plan;
  hipGraphExec_t exec = nullptr;
  uint64_t last_use = 0;
};

struct tmpm_session {
  tmpm_context* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::vector<SceneState> scenes;
  bool batch = false;
  int fixed64 = 0;
  bool valid = true;
  uint64_t pending = 0;
  bool prepared = false;

  HostWorld& world() { return scenes[0].world; }
  const HostWorld& world() const { return scenes[0].world; }

  // device state: enabled particles only (disabled ones never change and stay
  // on the host); SoA matrices use stride `cap`
  uint32_t n = 0;          // particle slots in use (a retired scene's slots stay until the next reorder)
  uint32_t n_enabled = 0;  // particles in device storage at the last upload
  uint32_t cap = 0;        // particle capacity of the buffers (SoA stride)
  uint32_t id_space = 0;   // insertion indices in use, all scenes, disabled particles included
  ParticleBuffers buf[2];
  int cur = 0;
  bool compacted_last = false;  // the last substep reordered: slots past the binned count are free
  // batches: state at the last synchronize, read per scene on demand
  ParticleBuffers synced;
  uint32_t synced_n = 0;
  DevBuf staging;
  DevBuf disabled_rec;     // snapshot records of disabled particles (2 float4 each)
  std::vector<float4> disabled_host;
  PackedParticles stage;                // admission staging, reused
  std::vector<HostWorld> spare_worlds;  // emptied worlds of retired scenes, reused
  uint32_t force_cap = 0;
  uint32_t last_pblocks = 0, last_gblocks = 0;  // block counts of the last synchronized substep
  uint32_t topo_cap = 0;   // particle capacity the topology's per-particle arrays were sized for
  std::map<ClassKey, uint32_t> class_ids;
  DevBuf force;            // float4 per particle, insertion order
  DevBuf classes, scene_plain, scene_force, obstacles, scene_err, planes;
  uint32_t obstacle_cap = 0, scene_cap = 0, plane_cap = 0;
  int shapes = 0;          // some obstacle is a triangle prism or convex polyhedron
  int cpic = 0;            // some scene uses CPIC obstacle groups
  bool plastic = false;    // some particle class is snow, sand or sand-NH: plastic state on the device
  // reactions and rigid coupling
  bool reactions_on = false;
  uint64_t react_saturated = 0;  // particle-side reaction impulses clamped to the bins' range, all substeps
  DevBuf react_bins, react_sub, react_total, rigid, rigid_maps, body_scene;
  uint32_t obstacle_total = 0, react_cap = 0, body_total = 0;
  bool react_active() const { return reactions_on || body_total > 0; }
  std::vector<DevClass> host_classes;
  uint32_t substeps_since_reorder = 1u << 30;
  // Substeps between moves of the particle state into block order: 16 was
  // the tuner's pick on every measured workload (1M and 4M scenes, batches,
  // 8k particles), 1.3% faster than 8 on the 1M scene.
  uint32_t reorder_interval = 16;

  // topology
  DevBuf owner, bkey, bucket_count, bucket_start, bucket_slot, bucket_pblock, grid_key, pblock_nbr,
      pblock_slot8, pblock_start, pblock_len, pslot, prank, sorted, counters, ptile, grid_rev, grid_vel, node_mask,
    bin_stage, bin_nstage, bin_claims, bin_nclaims, bin_csum, bmask, pmask;
  Topo topo{};
  uint32_t p2g_blocks = 4096, threads_pblock = 64, p2g_threads = 64;
  bool p2g_shape_fixed = false;  // P2G shape set by the tuner or an override, not by scene size
  bool g2p_shape_fixed = false;  // likewise for the G2P workgroup size
  int p2g_copies = 1;
  bool topo_valid = false;  // topology arrays describe the current device state
  bool table_dirty = true;  // block hash needs clearing before the next binning

  // FP64 audit (tmpm_session_set_audit): one record per scene slot and
  // substep, kept on the device for the pending substeps and copied to the
  // host at synchronize
  bool audit_on = false;
  DevBuf audit_dev;
  uint64_t audit_cap = 0;      // substep records the device buffer holds
  uint64_t audit_used = 0;     // records of the pending substeps
  uint32_t audit_stride = 0;   // scene slots per record set
  std::vector<double> audit_host;
  uint64_t audit_host_steps = 0;
  uint32_t audit_host_stride = 0;
  double* audit_next = nullptr;  // record set of the substep being enqueued

  // timing
  bool timing = false;
  // substep graphs (hipGraph): captured per distinct launch plan and replayed;
  // off by default (measured slower on the device for one substep per graph)
  bool use_graphs = false;
  std::vector<CachedGraph> graphs;
  uint64_t graph_clock = 0;
  std::vector<KernelTimer> pending_timers;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, uint64_t>> kernel_ms;

  // Bytes of every device allocation the session holds.
  uint64_t device_bytes() const {
    uint64_t total = 0;
    auto add = [&](const DevBuf& b) { total += b.ptr ? b.bytes : 0; };
    auto add_parts = [&](const ParticleBuffers& b) {
      for (const DevBuf* d : {&b.pos_mass, &b.vel_vol, &b.C, &b.F, &b.P, &b.cls, &b.ins_id}) add(*d);
    };
    add_parts(buf[0]);
    add_parts(buf[1]);
    add_parts(synced);
    for (const DevBuf* d :
         {&staging, &disabled_rec, &force, &classes, &scene_plain, &scene_force, &obstacles, &scene_err, &planes,
          &react_bins, &react_sub, &react_total, &rigid, &rigid_maps, &body_scene, &owner, &bkey, &bucket_count,
          &bucket_start, &bucket_slot, &bucket_pblock, &grid_key, &pblock_nbr, &pblock_slot8, &pblock_start,
          &pblock_len, &pslot, &prank, &sorted, &counters, &ptile, &grid_rev, &grid_vel, &node_mask, &bin_stage,
          &bin_nstage, &bin_claims, &bin_nclaims, &bin_csum, &bmask, &pmask, &audit_dev})
      add(*d);
    return total;
  }

  Parts parts(int which) const {
    Parts p{};
    const ParticleBuffers& b = buf[which];
    p.pos_mass = b.pos_mass.as<float4>();
    p.vel_vol = b.vel_vol.as<float4>();
    p.C = b.C.as<float>();
    p.F = b.F.as<float>();
    p.P = plastic ? b.P.as<float>() : nullptr;
    p.cls = b.cls.as<uint32_t>();
    p.force = force.as<float4>();
    p.ins_id = b.ins_id.as<uint32_t>();
    p.n = n;
    p.stride = cap;
    return p;
  }
};

// A world validated and packed for admission into a batch, not yet on any
// device (tmpm_batch_stage_scene / tmpm_batch_add_staged).
struct tmpm_staged_scene {
  SceneState scene;
  std::unique_ptr<PackedParticles> pk;
  std::vector<DevClass> class_table;  // distinct classes of the enabled particles, first seen first, scene index 0
  std::vector<uint32_t> class_local;  // per enabled particle: its class's index in class_table
  uint64_t pblocks = 0, gblocks = 0;
};

namespace {

// Kernel launch with optional event timing.
template <class F>
hipError_t timed(tmpm_session* s, const char* name, F&& launch) {
  if (!s->timing) {
    launch();
    return hipGetLastError();
  }
  KernelTimer t;
  t.name = name;
  for (hipEvent_t* e : {&t.start, &t.stop}) {
    if (!s->event_pool.empty()) {
      *e = s->event_pool.back();
      s->event_pool.pop_back();
    } else {
      (void)hipEventCreate(e);
    }
  }
  (void)hipEventRecord(t.start, s->stream);
  launch();
  (void)hipEventRecord(t.stop, s->stream);
  s->pending_timers.push_back(t);
  return hipGetLastError();
}

void collect_timers(tmpm_session* s) {
  for (auto& t : s->pending_timers) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      auto& e = s->kernel_ms[t.name];
      e.first += ms;
      e.second += 1;
    }
    s->event_pool.push_back(t.start);
    s->event_pool.push_back(t.stop);
  }
  s->pending_timers.clear();
}

uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)p;
}

// gpu.rs:664-677
tmpm_status validate_resident_world(SceneState& scene) {
  const HostWorld& w = scene.world;
  if (!params_valid(w.params)) return fail(TMPM_INVALID_INPUT, "invalid MPM parameters");
  const double dt = scene.substep_dt;
  if (!std::isfinite(dt) || dt <= 0.0 || dt > scene.stable_timestep())
    return fail(TMPM_INVALID_INPUT, "substep dt is not positive or exceeds the stable timestep");
  for (auto& o : w.obstacles)
    if (!obstacle_valid(o.o)) return fail(TMPM_INVALID_INPUT, "invalid obstacle");
  return TMPM_OK;
}

bool pack_obstacle(const tmpm_obstacle& o, DevObstacle& d) {
  std::memset(&d, 0, sizeof(d));
  d.shape = o.shape;
  d.boundary = o.boundary;
  bool ok = finite_f32(o.friction, d.friction);
  for (int k = 0; k < 3; ++k) {
    ok = finite_f32(o.center[k], d.center[k]) && ok;
    ok = finite_f32(o.linear_velocity[k], d.lin[k]) && ok;
    ok = finite_f32(o.angular_velocity[k], d.ang[k]) && ok;
  }
  for (int k = 0; k < 4; ++k) {
    ok = finite_f32(o.orientation[k], d.q[k]) && ok;
    ok = finite_f32(o.dims[k], d.dims[k]) && ok;
  }
  d.cpic_group = o.cpic_group;
  d.plane_begin = d.plane_count = 0;
  if (o.shape == TMPM_SHAPE_TRIANGLE_PRISM) {
    // gpu.rs pack_obstacle: f32 vertices must keep a non-degenerate triangle
    for (int k = 0; k < 9; ++k) ok = finite_f32(o.vertices[k], d.tri[k]) && ok;
    double e1[3], e2[3], cr[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)d.tri[3 + k] - d.tri[k];
      e2[k] = (double)d.tri[6 + k] - d.tri[k];
    }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double area2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    ok = ok && d.dims[0] > 0.0f && std::isfinite(area2) && area2 > 0.0;
    double r = 0.0;
    for (int v = 0; v < 3; ++v)
      r = std::max(r, std::sqrt(o.vertices[3 * v] * o.vertices[3 * v] + o.vertices[3 * v + 1] * o.vertices[3 * v + 1] +
                                o.vertices[3 * v + 2] * o.vertices[3 * v + 2]));
    ok = finite_f32((r + o.dims[0]) * (1.0 + 1e-5), d.bound) && ok;
  } else if (o.shape == TMPM_SHAPE_CONVEX) {
    ok = finite_f32(o.dims[0] * (1.0 + 1e-5), d.bound) && ok;
  }
  return ok;
}

// Convex planes of one obstacle in f32, appended to `out` (gpu.rs pack_obstacle).
tmpm_status pack_planes(const tmpm_obstacle& o, DevObstacle& d, std::vector<float4>& out) {
  if (o.plane_count > (uint32_t)kMaxConvexPlanes)
    return fail(TMPM_CAPACITY, "convex obstacle has more planes than the device supports (128)");
  d.plane_begin = (int32_t)out.size();
  d.plane_count = (int32_t)o.plane_count;
  for (uint32_t i = 0; i < o.plane_count; ++i) {
    float v[4];
    bool ok = true;
    for (int k = 0; k < 4; ++k) ok = finite_f32(o.planes[4 * i + k], v[k]) && ok;
    const float n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (!ok || !std::isfinite(n2) || n2 <= 0.0f) return fail(TMPM_INVALID_INPUT, "convex plane not representable in f32");
    out.push_back(make_float4(v[0], v[1], v[2], v[3]));
  }
  return TMPM_OK;
}

bool pack_rigid(const tmpm_rigid_body& b, DevRigid& d) {
  std::memset(&d, 0, sizeof(d));
  bool ok = finite_f32(b.inverse_mass, d.inv_mass) && b.inverse_mass >= 0.0;
  for (int k = 0; k < 3; ++k) {
    ok = finite_f32(b.position[k], d.pos[k]) && ok;
    ok = finite_f32(b.linear_velocity[k], d.lin[k]) && ok;
    ok = finite_f32(b.angular_velocity[k], d.ang[k]) && ok;
    ok = finite_f32(b.inverse_inertia[k], d.inv_inertia[k]) && b.inverse_inertia[k] >= 0.0 && ok;
  }
  double qn = 0.0;
  for (int k = 0; k < 4; ++k) {
    ok = finite_f32(b.orientation[k], d.q[k]) && ok;
    qn += b.orientation[k] * b.orientation[k];
  }
  return ok && std::fabs(qn - 1.0) < 1e-3;
}

// Rigid bodies of all stepping scenes, concatenated; each scene keeps its range.
tmpm_status upload_rigid(tmpm_session* s) {
  std::vector<DevRigid> packed;
  std::vector<uint32_t> owner;
  for (uint32_t k = 0; k < s->scenes.size(); ++k) {
    SceneState& scene = s->scenes[k];
    scene.dev.body_begin = (int32_t)packed.size();
    scene.dev.body_count = 0;
    scene.dev.rigid_integrate = 0;
    if (!scene.stepping()) continue;
    for (const auto& b : scene.bodies) {
      DevRigid d;
      if (!pack_rigid(b, d)) return fail(TMPM_INVALID_INPUT, "rigid body state not representable");
      packed.push_back(d);
      owner.push_back(k);
    }
    scene.dev.body_count = (int32_t)scene.bodies.size();
    scene.dev.rigid_integrate = scene.rigid_integrate ? 1 : 0;
    for (int a = 0; a < 3; ++a)
      if (!finite_f32(scene.rigid_gravity[a], scene.dev.rigid_gravity[a]))
        return fail(TMPM_INVALID_INPUT, "rigid gravity not representable");
    if (!finite_f32(scene.rigid_max_speed, scene.dev.rigid_max_speed) || scene.rigid_max_speed < 0.0)
      return fail(TMPM_INVALID_INPUT, "rigid speed cap invalid");
  }
  s->body_total = (uint32_t)packed.size();
  if (packed.empty()) return TMPM_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(s->rigid.alloc(sizeof(DevRigid) * packed.size()));
  HIP_TRY(s->body_scene.alloc(4 * owner.size()));
  HIP_TRY(hipMemcpy(s->rigid.ptr, packed.data(), sizeof(DevRigid) * packed.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->body_scene.ptr, owner.data(), 4 * owner.size(), hipMemcpyHostToDevice));
  return TMPM_OK;
}

// Largest distance from an obstacle's center to a point within one cell of
// it: the lever arm that bounds its angular reaction contributions.
double obstacle_reach(const tmpm_obstacle& o, double h) {
  double r = 0.0;
  for (int k = 0; k < 4; ++k) r += std::fabs(o.dims[k]);
  for (int k = 0; k < 9; ++k) r = std::max(r, std::fabs(o.vertices[k]));
  return 2.0 * r + 2.0 * h;
}

// Obstacles of all active scenes, concatenated; each scene keeps its range.
tmpm_status upload_obstacles(tmpm_session* s) {
  std::vector<DevObstacle> packed;
  std::vector<DevMapping> maps;
  std::vector<float4> planes;
  bool shapes = false, cpic = false;
  for (auto& scene : s->scenes) {
    scene.dev.obstacle_begin = (int32_t)packed.size();
    scene.dev.obstacle_count = 0;
    if (!scene.stepping()) continue;
    const bool mapped = !scene.bodies.empty();
    if (mapped && scene.maps.size() != scene.world.obstacles.size())
      return fail(TMPM_INVALID_INPUT, "one rigid mapping per obstacle is required");
    for (auto& ob : scene.world.obstacles) {
      const tmpm_obstacle& o = ob.o;
      DevObstacle d;
      if (!pack_obstacle(o, d)) return fail(TMPM_INVALID_INPUT, "obstacle is not representable in f32");
      // gpu.rs cpic_obstacles: groups count only while the world uses CPIC
      if (!scene.world.params.use_cpic) d.cpic_group = -1;
      cpic = cpic || d.cpic_group >= 0;
      if (o.shape == TMPM_SHAPE_CONVEX) {
        tmpm_status st = pack_planes(o, d, planes);
        if (st != TMPM_OK) return st;
      }
      shapes = shapes || o.shape == TMPM_SHAPE_TRIANGLE_PRISM || o.shape == TMPM_SHAPE_CONVEX;
      // reaction units: momentum at the scene's P2G scale; angular impulse
      // scaled down by the lever-arm bound so both stay well inside int64
      const int lin = std::max(-90, std::min(150, scene.dev.mom_shift));
      const int ang = std::max(-90, std::min(150, lin - (int)std::ceil(std::log2(obstacle_reach(o, scene.world.params.cell_width)))));
      d.rshift_lin = lin;
      d.rshift_ang = ang;
      d.rscale_lin = std::ldexp(1.0f, lin - 32);
      d.rscale_ang = std::ldexp(1.0f, ang - 32);
      DevMapping m{};
      m.body = -1;
      if (mapped) {
        const tmpm_rigid_mapping& src = scene.maps[packed.size() - scene.dev.obstacle_begin];
        if (src.body != UINT32_MAX) {
          if (src.body >= scene.bodies.size()) return fail(TMPM_INVALID_INPUT, "rigid mapping names no body");
          m.body = scene.dev.body_begin + (int32_t)src.body;
        }
        bool ok = true;
        for (int k = 0; k < 3; ++k) ok = finite_f32(src.local_center[k], m.local_center[k]) && ok;
        for (int k = 0; k < 4; ++k) ok = finite_f32(src.local_orientation[k], m.local_q[k]) && ok;
        if (!ok) return fail(TMPM_INVALID_INPUT, "rigid mapping not representable");
      }
      packed.push_back(d);
      maps.push_back(m);
    }
    scene.dev.obstacle_count = (int32_t)(packed.size() - scene.dev.obstacle_begin);
    if (scene.dev.obstacle_count > kMaxObstacles) return fail(TMPM_CAPACITY, "too many obstacles");
  }
  if (packed.size() > s->obstacle_cap || !s->obstacles.ptr) {
    s->obstacle_cap = std::max<uint32_t>(16, (uint32_t)packed.size());
    HIP_TRY(s->obstacles.alloc(sizeof(DevObstacle) * s->obstacle_cap));
    HIP_TRY(s->rigid_maps.alloc(sizeof(DevMapping) * s->obstacle_cap));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (!packed.empty()) {
    HIP_TRY(hipMemcpy(s->obstacles.ptr, packed.data(), packed.size() * sizeof(DevObstacle), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->rigid_maps.ptr, maps.data(), maps.size() * sizeof(DevMapping), hipMemcpyHostToDevice));
  }
  if (planes.size() > s->plane_cap || !s->planes.ptr) {
    s->plane_cap = std::max<uint32_t>(16, (uint32_t)planes.size());
    HIP_TRY(s->planes.alloc(sizeof(float4) * s->plane_cap));
  }
  if (!planes.empty())
    HIP_TRY(hipMemcpy(s->planes.ptr, planes.data(), planes.size() * sizeof(float4), hipMemcpyHostToDevice));
  s->topo.planes = s->planes.as<float4>();
  s->shapes = shapes ? 1 : 0;
  s->cpic = cpic ? 1 : 0;
  s->obstacle_total = (uint32_t)packed.size();
  s->topo.obstacle_total = s->obstacle_total;
  if (s->react_active() && s->obstacle_total > 0) {
    if (s->obstacle_total > s->react_cap || !s->react_bins.ptr) {
      s->react_cap = std::max<uint32_t>(16, s->obstacle_total);
      HIP_TRY(s->react_bins.alloc(8ull * kReactBins * 6 * s->react_cap));
      HIP_TRY(s->react_sub.alloc(4ull * 6 * s->react_cap));
      HIP_TRY(s->react_total.alloc(8ull * 6 * s->react_cap));
    }
    HIP_TRY(hipMemset(s->react_bins.ptr, 0, 8ull * kReactBins * 6 * s->react_cap));
    HIP_TRY(hipMemset(s->react_sub.ptr, 0, 4ull * 6 * s->react_cap));
    HIP_TRY(hipMemset(s->react_total.ptr, 0, 8ull * 6 * s->react_cap));
    s->topo.react_bins = s->react_bins.as<long long>();
    s->topo.react_sub = s->react_sub.as<float>();
    s->topo.react_total = s->react_total.as<double>();
  }
  return TMPM_OK;
}

// Scene tables: `plain` for ordinary substeps, `force` (apply_force set for
// scenes with a one-shot force edit) for the first substep after an edit.
tmpm_status upload_scene(tmpm_session* s) {
  const size_t n = s->scenes.size();
  if (n > s->scene_cap || !s->scene_plain.ptr) {
    s->scene_cap = (uint32_t)std::max<size_t>(n, 1);
    HIP_TRY(s->scene_plain.alloc(sizeof(DevScene) * s->scene_cap));
    HIP_TRY(s->scene_force.alloc(sizeof(DevScene) * s->scene_cap));
    HIP_TRY(s->scene_err.alloc(4ull * s->scene_cap));
    HIP_TRY(hipMemset(s->scene_err.ptr, 0, 4ull * s->scene_cap));
  }
  std::vector<DevScene> plain(n), forced(n);
  for (size_t k = 0; k < n; ++k) {
    plain[k] = forced[k] = s->scenes[k].dev;
    plain[k].apply_force = 0;
    forced[k].apply_force = s->scenes[k].force_pending ? 1 : 0;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(s->scene_plain.ptr, plain.data(), sizeof(DevScene) * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->scene_force.ptr, forced.data(), sizeof(DevScene) * n, hipMemcpyHostToDevice));
  s->topo.scene_err = s->scene_err.as<uint32_t>();
  s->topo.isolate = s->batch ? 1u : 0u;
  return TMPM_OK;
}

// Scene constants; f32 packing as in gpu.rs:1118-1127, 1407-1432.
// Packs scene k's parameters (keeps its obstacle range).
tmpm_status build_scene(SceneState& scene) {
  const tmpm_params& p = scene.world.params;
  DevScene& sc = scene.dev;
  float h, step, inv_d;
  if (!finite_f32(p.cell_width, h) || !finite_f32(scene.substep_dt, step) ||
      !finite_f32(4.0 / (p.cell_width * p.cell_width), inv_d) || h <= 0.0f || step <= 0.0f)
    return fail(TMPM_INVALID_INPUT, "cell width or dt not representable in f32");
  float inv_h = 1.0f / h;
  if (!std::isfinite(inv_h)) return fail(TMPM_INVALID_INPUT, "cell width too small");
  sc.h = h;
  sc.inv_h = inv_h;
  sc.inv_d = inv_d;
  sc.dt = step;
  for (int k = 0; k < 3; ++k)
    if (!finite_f32(p.gravity[k], sc.gravity[k])) return fail(TMPM_INVALID_INPUT, "gravity not representable");
  sc.has_bounds = p.has_bounds ? 1 : 0;
  // Fixed-point exponents: a particle's mass maps to at most 2^43 units and
  // every contribution must stay below 2^45 units (checked per particle);
  // momentum is scaled for mass x 64 cells per substep.  The unit is fine
  // enough that P2G's rounding to the nearest unit stays far below f32
  // resolution.
  if (scene.max_mass > 0.0) {
    double vbound = 64.0 * p.cell_width / scene.substep_dt;
    const int top = 43;
    int ms = (int)std::floor(std::log2(std::ldexp(1.0, top) / scene.max_mass));
    int ps = (int)std::floor(std::log2(std::ldexp(1.0, top) / (scene.max_mass * vbound)));
    sc.mass_shift = std::max(-120, std::min(120, ms));
    sc.mom_shift = std::max(-120, std::min(120, ps));
  }
  for (int a = 0; a < 3; ++a) {
    sc.first_base[a] = std::numeric_limits<int32_t>::min();
    sc.last_base[a] = 2147483645;
    sc.bmin[a] = sc.bmax[a] = 0.0f;
    if (p.has_bounds) {
      if (!finite_f32(p.bound_min[a], sc.bmin[a]) || !finite_f32(p.bound_max[a], sc.bmax[a]) ||
          sc.bmin[a] >= sc.bmax[a])
        return fail(TMPM_INVALID_INPUT, "bounds not representable in f32");
      // gpu.wgsl:962-964: resident particles must stay within two cells of the bounds.
      volatile float lo = sc.bmin[a] * inv_h;
      volatile float hi = sc.bmax[a] * inv_h;
      volatile float lo2 = lo - 0.5f, hi2 = hi - 0.5f;
      sc.first_base[a] = (int32_t)std::floor((float)lo2) - 2;
      sc.last_base[a] = (int32_t)std::floor((float)hi2) + 2;
    }
  }
  return TMPM_OK;
}

// Launch-shape overrides for tuning sweeps, read once per session.
void read_tuning_env(tmpm_session* s) {
  if (const char* v = std::getenv("TMPM_PBLOCK_THREADS")) {
    int th = std::atoi(v);
    if (th == 64 || th == 128 || th == 192 || th == 256) s->threads_pblock = (uint32_t)th;
    s->g2p_shape_fixed = true;
  }
  if (const char* v = std::getenv("TMPM_P2G_THREADS")) {
    int th = std::atoi(v);
    if (th == 64 || th == 128) s->p2g_threads = (uint32_t)th;
    s->p2g_shape_fixed = true;
  }
  if (const char* v = std::getenv("TMPM_P2G_COPIES")) {
    s->p2g_copies = std::atoi(v) == 1 ? 1 : 2;
    s->p2g_shape_fixed = true;
  }
  if (const char* v = std::getenv("TMPM_GRAPHS")) s->use_graphs = std::atoi(v) != 0;
  if (const char* v = std::getenv("TMPM_REORDER_INTERVAL")) {
    int r = std::atoi(v);
    if (r >= 1 && r <= 1000000) s->reorder_interval = (uint32_t)r;
  }
}

// Allocates the topology for `pblocks` particle blocks and `gblocks` grid
// blocks of capacity.
tmpm_status alloc_topology(tmpm_session* s, uint32_t pblock_cap, uint32_t grid_cap) {
  pblock_cap = std::min(pblock_cap, grid_cap);  // a particle block's grid slot is its id
  uint32_t buckets = std::max<uint32_t>(next_pow2(2ull * grid_cap), (uint32_t)kMinBuckets);
  if ((uint64_t)s->cap + 8ull * buckets + 1 >= 0xffffffffull) return fail(TMPM_CAPACITY, "topology id space exhausted");
  Topo& t = s->topo;
  HIP_TRY(s->owner.alloc(4ull * buckets));
  HIP_TRY(s->bkey.alloc(16ull * buckets));
  HIP_TRY(s->bmask.alloc(4ull * buckets));
  HIP_TRY(s->bucket_count.alloc(4ull * buckets));
  HIP_TRY(s->bucket_start.alloc(4ull * buckets));
  HIP_TRY(s->bucket_slot.alloc(4ull * buckets));
  HIP_TRY(s->bucket_pblock.alloc(4ull * buckets));
  HIP_TRY(s->grid_key.alloc(16ull * grid_cap));
  HIP_TRY(s->pblock_nbr.alloc(32ull * pblock_cap));
  HIP_TRY(s->pblock_slot8.alloc(32ull * pblock_cap));
  HIP_TRY(s->pblock_start.alloc(4ull * pblock_cap));
  HIP_TRY(s->pblock_len.alloc(4ull * pblock_cap));
  s->topo_cap = s->cap;
  HIP_TRY(s->pslot.alloc(4ull * std::max<uint32_t>(s->cap, 1)));
  HIP_TRY(s->prank.alloc(4ull * std::max<uint32_t>(s->cap, 1)));
  HIP_TRY(s->sorted.alloc(4ull * std::max<uint32_t>(s->cap, 1)));
  HIP_TRY(s->pmask.alloc(4ull * std::max<uint32_t>(s->cap, 1)));
  HIP_TRY(s->counters.alloc(4ull * CNT_COUNT));
  HIP_TRY(hipMemset(s->counters.ptr, 0, 4ull * CNT_COUNT));
  // Binning workgroups: at most ~2048 (each one's id prefix sums over all of
  // them), each over a whole number of 256-particle rounds.
  const uint64_t nparts = std::max<uint32_t>(s->cap, 1);
  const uint32_t rounds = (uint32_t)((nparts + 256ull * 2048 - 1) / (256ull * 2048));
  t.bin_span = 256u * rounds;
  t.bin_groups = (uint32_t)((nparts + t.bin_span - 1) / t.bin_span);
  const uint64_t staged = (uint64_t)t.bin_groups * t.bin_span;
  HIP_TRY(s->bin_stage.alloc(4ull * staged));
  HIP_TRY(s->bin_nstage.alloc(28ull * staged));
  HIP_TRY(s->bin_claims.alloc(4ull * t.bin_groups));
  HIP_TRY(s->bin_nclaims.alloc(4ull * t.bin_groups));
  HIP_TRY(s->bin_csum.alloc(4ull * t.bin_groups));
  t.bin_stage = s->bin_stage.as<uint32_t>();
  t.bin_nstage = s->bin_nstage.as<uint32_t>();
  t.bin_claims = s->bin_claims.as<uint32_t>();
  t.bin_nclaims = s->bin_nclaims.as<uint32_t>();
  t.bin_csum = s->bin_csum.as<uint32_t>();
  HIP_TRY(s->ptile.alloc((size_t)grid_cap * kTileNodes * (s->fixed64 ? 32 : 16)));
  HIP_TRY(s->grid_rev.alloc(32ull * grid_cap));
  HIP_TRY(s->grid_vel.alloc((size_t)grid_cap * kBlockNodes * 16));
  HIP_TRY(s->node_mask.alloc(8ull * grid_cap));
  t.owner = s->owner.as<uint32_t>();
  t.bkey = s->bkey.as<int4>();
  t.bmask = s->bmask.as<uint32_t>();
  t.pmask = s->pmask.as<uint32_t>();
  t.bucket_count = s->bucket_count.as<uint32_t>();
  t.bucket_start = s->bucket_start.as<uint32_t>();
  t.bucket_slot = s->bucket_slot.as<uint32_t>();
  t.bucket_pblock = s->bucket_pblock.as<uint32_t>();
  t.bucket_mask = buckets - 1;
  t.grid_key = s->grid_key.as<int4>();
  t.grid_cap = grid_cap;
  t.pblock_cap = pblock_cap;
  t.pblock_nbr = s->pblock_nbr.as<uint32_t>();
  t.pblock_slot8 = s->pblock_slot8.as<uint32_t>();
  t.pblock_start = s->pblock_start.as<uint32_t>();
  t.pblock_len = s->pblock_len.as<uint32_t>();
  t.pslot = s->pslot.as<uint32_t>();
  t.prank = s->prank.as<uint32_t>();
  t.sorted = s->sorted.as<uint32_t>();
  t.counters = s->counters.as<uint32_t>();
  t.ptile = s->ptile.ptr;
  t.grid_rev = s->grid_rev.as<uint32_t>();
  t.grid_vel = s->grid_vel.as<float4>();
  t.node_mask = s->node_mask.as<uint32_t>();
  s->p2g_blocks = std::min<uint32_t>(pblock_cap, 65535);
  s->topo_valid = false;
  s->table_dirty = true;
  return TMPM_OK;
}

// Counts particle blocks and grid blocks (block + its 7 upper neighbours) of
// the enabled particles on the host, f32 cell rule.
template <class FV, class UV>
void count_blocks(const FV& pos32, const UV& region, const FV& inv_h, uint64_t& pblocks, uint64_t& gblocks) {
  // Dense path for one region: occupancy bytes over the bounding box of the
  // particle blocks (grid blocks are those dilated by one toward +x, +y, +z).
  // Chunks of 8192 particles keep every host thread busy for scenes of a few
  // hundred thousand particles (an admission's block count).
  constexpr size_t kCountChunk = 8192;
  const size_t np = region.size();
  pblocks = gblocks = 0;
  if (np == 0) return;
  bool one_region = true;
  for (size_t i = 1; i < np && one_region; ++i) one_region = region[i] == region[0];
  if (one_region) {
    const size_t nchunk = (np + kCountChunk - 1) / kCountChunk;
    std::vector<std::array<int32_t, 6>> box(nchunk, {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN});
    parallel_for(nchunk, [&](size_t cb, size_t ce) {
      for (size_t c = cb; c < ce; ++c)
        for (size_t i = c * kCountChunk; i < std::min(np, (c + 1) * kCountChunk); ++i) {
          int32_t cc[3];
          bool ok = true;
          for (int a = 0; a < 3; ++a) ok = cell_f32(pos32[3 * i + a], inv_h[i], cc[a]) && ok;
          if (!ok) continue;
          for (int a = 0; a < 3; ++a) {
            box[c][a] = std::min(box[c][a], cc[a] >> 2);
            box[c][3 + a] = std::max(box[c][3 + a], cc[a] >> 2);
          }
        }
    }, 1);
    std::array<int64_t, 6> bb{INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    for (auto& b : box)
      for (int a = 0; a < 3; ++a) {
        bb[a] = std::min<int64_t>(bb[a], b[a]);
        bb[3 + a] = std::max<int64_t>(bb[3 + a], b[3 + a]);
      }
    if (bb[0] > bb[3]) return;  // no representable particle
    const int64_t ex = bb[3] - bb[0] + 2, ey = bb[4] - bb[1] + 2, ez = bb[5] - bb[2] + 2;
    if (ex * ey * ez <= (1 << 24)) {
      std::unique_ptr<std::atomic<uint8_t>[]> occ(new std::atomic<uint8_t>[ex * ey * ez]());
      parallel_for(nchunk, [&](size_t cb, size_t ce) {
        for (size_t c = cb; c < ce; ++c)
          for (size_t i = c * kCountChunk; i < std::min(np, (c + 1) * kCountChunk); ++i) {
            int32_t cc[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) ok = cell_f32(pos32[3 * i + a], inv_h[i], cc[a]) && ok;
            if (!ok) continue;
            const int64_t x = (cc[0] >> 2) - bb[0], y = (cc[1] >> 2) - bb[1], z = (cc[2] >> 2) - bb[2];
            occ[(x * ey + y) * ez + z].store(1, std::memory_order_relaxed);
          }
      }, 1);
      for (int64_t x = 0; x < ex; ++x)
        for (int64_t y = 0; y < ey; ++y)
          for (int64_t z = 0; z < ez; ++z) {
            auto at = [&](int64_t i, int64_t j, int64_t k) {
              return i >= 0 && j >= 0 && k >= 0 && occ[(i * ey + j) * ez + k].load(std::memory_order_relaxed) != 0;
            };
            pblocks += at(x, y, z);
            bool g = false;
            for (int d = 0; d < 8 && !g; ++d) g = at(x - (d >> 2), y - ((d >> 1) & 1), z - (d & 1));
            gblocks += g;
          }
      return;
    }
  }

  struct K {
    int32_t x, y, z;
    uint32_t r;
    bool operator==(const K& o) const { return x == o.x && y == o.y && z == o.z && r == o.r; }
  };
  struct H {
    size_t operator()(const K& k) const {
      return ((size_t)(uint32_t)k.x * 73856093u) ^ ((size_t)(uint32_t)k.y * 19349663u) ^
             ((size_t)(uint32_t)k.z * 83492791u) ^ ((size_t)k.r * 2654435761u);
    }
  };
  std::unordered_map<K, char, H> pb, gb;
  size_t n = region.size();
  K last{0, 0, 0, 0};
  bool have_last = false;
  for (size_t i = 0; i < n; ++i) {
    int32_t c[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) ok = cell_f32(pos32[3 * i + a], inv_h[i], c[a]) && ok;
    if (!ok) continue;
    K k{c[0] >> 2, c[1] >> 2, c[2] >> 2, region[i]};
    if (have_last && k == last) continue;  // neighbours in insertion order usually share a block
    last = k;
    have_last = true;
    if (pb.emplace(k, 1).second)
      for (int d = 0; d < 8; ++d) gb.emplace(K{k.x + (d >> 2), k.y + ((d >> 1) & 1), k.z + (d & 1), k.r}, 1);
  }
  pblocks = pb.size();
  gblocks = gb.size();
}

// Packs a particle's class (material constants, radius, damping, fixed flag,
// region, the scene's cell scale).  `scene` goes into the region's high 16 bits.
tmpm_status pack_class(const HostWorld& w, size_t i, uint32_t scene, float inv_h, float rad, float damp,
                       DevClass& c) {
  const tmpm_material& m = w.material[i];
  c = DevClass{};
  c.region = (scene << 16) | w.color[i];
  c.inv_h = inv_h;
  c.fixed = w.fixed[i] ? 1u : 0u;
  c.radius = rad;
  c.damping = damp;
  if (!w.enabled[i]) return TMPM_OK;
  switch (m.kind) {
    case TMPM_MATERIAL_LINEAR_ELASTIC:
    case TMPM_MATERIAL_NEO_HOOKEAN: {
      double lambda, mu;
      lame(m.p[0], m.p[1], lambda, mu);
      if (!finite_f32(lambda, c.lambda) || !finite_f32(mu, c.mu))
        return fail(TMPM_INVALID_INPUT, "Lame parameters not representable");
      c.kind = m.kind == TMPM_MATERIAL_LINEAR_ELASTIC ? DM_COROTATED : DM_NEO_HOOKEAN;
      c.wave_numerator = c.lambda + 2.0f * c.mu;
      return TMPM_OK;
    }
    case TMPM_MATERIAL_FLUID:
      c.kind = DM_FLUID;
      if (!finite_f32(m.p[0], c.bulk) || !finite_f32(m.p[1], c.gamma) || !finite_f32(m.p[2], c.viscosity) ||
          !finite_f32(m.p[3], c.tensile))
        return fail(TMPM_INVALID_INPUT, "fluid parameters not representable");
      c.wave_numerator = c.bulk * c.gamma;
      return TMPM_OK;
    case TMPM_MATERIAL_SNOW:
    case TMPM_MATERIAL_SAND:
    case TMPM_MATERIAL_SAND_NEO_HOOKEAN: {
      double lambda, mu;
      lame(m.p[0], m.p[1], lambda, mu);
      if (!finite_f32(lambda, c.lambda) || !finite_f32(mu, c.mu))
        return fail(TMPM_INVALID_INPUT, "Lame parameters not representable");
      c.wave_numerator = c.lambda + 2.0f * c.mu;
      bool ok = true;
      if (m.kind == TMPM_MATERIAL_SNOW) {
        c.kind = DM_SNOW;
        ok = finite_f32(m.p[2], c.crit_compression) && finite_f32(m.p[3], c.crit_stretch) &&
             finite_f32(m.p[4], c.hardening);
      } else {
        c.kind = m.kind == TMPM_MATERIAL_SAND ? DM_SAND : DM_SAND_NH;
        ok = finite_f32(m.p[2], c.friction_angle) && finite_f32(m.p[3], c.cohesion);
      }
      if (!ok) return fail(TMPM_INVALID_INPUT, "plastic material parameters not representable");
      return TMPM_OK;
    }
    default:
      return fail(TMPM_INVALID_INPUT, "unknown material");
  }
}


uint32_t intern_class(tmpm_session* s, const DevClass& c) {
  // runs of particles usually share a class: compare with the last one first
  static thread_local const tmpm_session* last_s = nullptr;
  static thread_local uint32_t last_id = 0;
  if (last_s == s && last_id < s->host_classes.size() &&
      std::memcmp(&s->host_classes[last_id], &c, sizeof(DevClass)) == 0)
    return last_id;
  auto key = class_key(c);
  auto it = s->class_ids.find(key);
  uint32_t id;
  if (it != s->class_ids.end()) {
    id = it->second;
  } else {
    id = (uint32_t)s->host_classes.size();
    s->class_ids.emplace(key, id);
    s->host_classes.push_back(c);
  }
  last_s = s;
  last_id = id;
  return id;
}

// Validates and converts particle i of `w`; *en tells whether it is enabled.
tmpm_status pack_one(const HostWorld& w, size_t i, uint32_t k, float inv_h, float f[8], float cf[9], float ff[9],
                     float pp[3], DevClass& c, bool* en) {
  *en = w.enabled[i] != 0;
  bool ok = true;
  for (int a = 0; a < 3; ++a) ok = finite_f32(w.pos[3 * i + a], f[a]) && ok;
  ok = finite_f32(*en ? w.mass[i] : 0.0, f[3]) && ok;
  for (int a = 0; a < 3; ++a) ok = finite_f32(w.vel[3 * i + a], f[4 + a]) && ok;
  ok = finite_f32(w.rest_volume[i], f[7]) && ok;
  float rad, damp;
  ok = finite_f32(w.radius[i], rad) && finite_f32(w.damping[i], damp) && ok;
  if (!ok) return fail(TMPM_INVALID_INPUT, "particle state not representable in f32");
  if (*en && (f[3] <= 0.0f || f[7] <= 0.0f))
    return fail(TMPM_INVALID_INPUT, "enabled particle mass or volume underflows f32");
  if (rad <= 0.0f || w.damping[i] < 0.0) return fail(TMPM_INVALID_INPUT, "radius or damping invalid");
  if (!*en) return TMPM_OK;
  for (int q = 0; q < 9; ++q)
    if (!finite_f32(w.C[9 * i + q], cf[q]) || !finite_f32(w.F[9 * i + q], ff[q]))
      return fail(TMPM_INVALID_INPUT, "particle matrices not representable in f32");
  for (int q = 0; q < 3; ++q)
    if (!finite_f32(w.plastic[3 * i + q], pp[q])) return fail(TMPM_INVALID_INPUT, "plastic state not representable in f32");
  tmpm_status st = pack_class(w, i, k, inv_h, rad, damp, c);
  if (st != TMPM_OK) return st;
  // base cell must be representable (gpu.rs:1143-1153)
  for (int a = 0; a < 3; ++a) {
    int32_t cc;
    if (!cell_f32(f[a], inv_h, cc)) return fail(TMPM_INVALID_INPUT, "particle cell not representable");
  }
  return TMPM_OK;
}

// Whether any interned particle class carries a plastic state.
bool classes_plastic(const tmpm_session* s) {
  for (const DevClass& c : s->host_classes)
    if (dm_plastic(c.kind)) return true;
  return false;
}

// The host part of pack_scene: every enabled particle packed with scene
// index k and insertion base scene.first, and its (not yet interned) class.
// Two parallel passes over fixed chunks (validate and count, then fill at
// the chunks' offsets); the error reported is the one of the lowest failing
// index, as in one pass.
tmpm_status pack_scene_host(const SceneState& scene, uint32_t k, PackedParticles& out, std::vector<DevClass>& classes) {
  const HostWorld& w = scene.world;
  const float inv_h = scene.dev.inv_h;
  const size_t n = w.size();
  constexpr size_t kChunk = 8192;
  const size_t chunks = (n + kChunk - 1) / kChunk;
  struct ChunkInfo {
    size_t enabled = 0, disabled = 0;
    size_t bad = SIZE_MAX;
    tmpm_status st = TMPM_OK;
    std::string msg;
  };
  std::vector<ChunkInfo> info(chunks);
  classes.resize(n);
  parallel_for(chunks, [&](size_t cb, size_t ce) {
    for (size_t c = cb; c < ce; ++c) {
      ChunkInfo& ci = info[c];
      for (size_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); ++i) {
        float f[8], cf[9], ff[9], pp[3];
        bool en;
        tmpm_status st = pack_one(w, i, k, inv_h, f, cf, ff, pp, classes[i], &en);
        if (st != TMPM_OK) {
          ci.bad = i;
          ci.st = st;
          ci.msg = g_last_error;
          break;
        }
        (en ? ci.enabled : ci.disabled)++;
      }
    }
  }, 1);
  size_t ne = 0, nd = 0;
  std::vector<size_t> eoff(chunks), doff(chunks);
  for (size_t c = 0; c < chunks; ++c) {
    if (info[c].bad != SIZE_MAX) return fail(info[c].st, info[c].msg);
    eoff[c] = ne;
    doff[c] = nd;
    ne += info[c].enabled;
    nd += info[c].disabled;
  }
  const size_t e0 = out.pm.size(), d0 = out.disabled.size();
  out.pm.resize(e0 + ne);
  out.vv.resize(e0 + ne);
  out.C.resize(9 * (e0 + ne));
  out.F.resize(9 * (e0 + ne));
  out.Pl.resize(3 * (e0 + ne));
  out.cls.resize(e0 + ne);
  out.ins.resize(e0 + ne);
  out.pos32.resize(3 * (e0 + ne));
  out.inv_h.resize(e0 + ne, inv_h);
  out.region.resize(e0 + ne);
  out.disabled.resize(d0 + 2 * nd);
  parallel_for(chunks, [&](size_t cb, size_t ce) {
    for (size_t c = cb; c < ce; ++c) {
      size_t e = e0 + eoff[c], d = d0 + 2 * doff[c];
      for (size_t i = c * kChunk; i < std::min(n, (c + 1) * kChunk); ++i) {
        float f[8], cf[9], ff[9], pp[3];
        bool en;
        DevClass dc;
        pack_one(w, i, k, inv_h, f, cf, ff, pp, dc, &en);  // validated in the first pass
        const uint32_t g = scene.first + (uint32_t)i;
        if (!en) {
          out.disabled[d++] = make_float4(f[0], f[1], f[2], __builtin_bit_cast(float, g));
          out.disabled[d++] = make_float4(f[4], f[5], f[6], f[7]);
          continue;
        }
        for (int a = 0; a < 3; ++a) out.pos32[3 * e + a] = f[a];
        out.region[e] = classes[i].region;
        out.pm[e] = make_float4(f[0], f[1], f[2], f[3]);
        out.vv[e] = make_float4(f[4], f[5], f[6], f[7]);
        std::memcpy(&out.C[9 * e], cf, sizeof cf);
        std::memcpy(&out.F[9 * e], ff, sizeof ff);
        std::memcpy(&out.Pl[3 * e], pp, sizeof pp);
        out.ins[e] = g;
        ++e;
      }
    }
  }, 1);
  return TMPM_OK;
}

// Interns the classes of a packed scene's enabled particles into out.cls
// from index e0 on: sequential, ids in first-seen order; runs share a class.
void intern_scene_classes(tmpm_session* s, const HostWorld& w, const std::vector<DevClass>& classes,
                          PackedParticles& out, size_t e0) {
  size_t e = e0;
  for (size_t i = 0; i < w.size(); ++i)
    if (w.enabled[i]) out.cls[e++] = intern_class(s, classes[i]);
}

// Appends scene k's enabled particles to `out` (insertion index first + i)
// and records its disabled ones for snapshots.
tmpm_status pack_scene(tmpm_session* s, uint32_t k, PackedParticles& out) {
  const size_t e0 = out.pm.size();
  std::vector<DevClass> classes;
  tmpm_status st = pack_scene_host(s->scenes[k], k, out, classes);
  if (st != TMPM_OK) return st;
  intern_scene_classes(s, s->scenes[k].world, classes, out, e0);
  return TMPM_OK;
}

tmpm_status alloc_particles(ParticleBuffers& pb, uint32_t cap) {
  HIP_TRY(pb.pos_mass.alloc(16ull * cap));
  HIP_TRY(pb.vel_vol.alloc(16ull * cap));
  HIP_TRY(pb.C.alloc(36ull * cap));
  HIP_TRY(pb.F.alloc(36ull * cap));
  HIP_TRY(pb.P.alloc(12ull * cap));
  HIP_TRY(pb.cls.alloc(4ull * cap));
  HIP_TRY(pb.ins_id.alloc(4ull * cap));
  return TMPM_OK;
}

// Copies the first n slots of `src` (stride src_cap) into `dst` (stride dst_cap).
tmpm_status copy_particles(ParticleBuffers& dst, uint32_t dst_cap, const ParticleBuffers& src, uint32_t src_cap,
                           uint32_t n, hipStream_t st) {
  if (n == 0) return TMPM_OK;
  auto view = [](const ParticleBuffers& b, uint32_t cap) {
    Parts v{};
    v.pos_mass = b.pos_mass.as<float4>();
    v.vel_vol = b.vel_vol.as<float4>();
    v.C = b.C.as<float>();
    v.F = b.F.as<float>();
    v.P = b.P.as<float>();
    v.cls = b.cls.as<uint32_t>();
    v.ins_id = b.ins_id.as<uint32_t>();
    v.stride = cap;
    return v;
  };
  launch_copy_state(view(dst, dst_cap), view(src, src_cap), n, st);
  HIP_TRY(hipGetLastError());
  return TMPM_OK;
}

// Writes packed particles into the current buffer at slots [at, at + size).
// Component-major copies of C, F and the plastic state (9 + 9 + 3 columns of
// n floats), the layout of the device buffers, in the packing's page-locked
// staging.  A staged admission builds them on its host thread, so that the
// upload between ticks is copies only.
void pack_columns(PackedParticles& pk) {
  const size_t n = pk.size();
  auto& cols = pk.cols;
  cols.resize(21 * n);
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      for (int q = 0; q < 9; ++q) {
        cols[(size_t)q * n + i] = pk.C[9 * i + q];
        cols[(size_t)(9 + q) * n + i] = pk.F[9 * i + q];
      }
      for (int q = 0; q < 3; ++q) cols[(size_t)(18 + q) * n + i] = pk.Pl[3 * i + q];
    }
  });
}

// Uploads the packed particles into slots [at, at + n) of the live buffers:
// copies on the session stream from page-locked staging, one wait at the end
// (the staging may be reused once this returns).
tmpm_status upload_particles(tmpm_session* s, PackedParticles& pk, uint32_t at) {
  const size_t n = pk.size();
  if (n == 0) return TMPM_OK;
  ParticleBuffers& pb = s->buf[s->cur];
  if (pk.cols.size() != 21 * n) pack_columns(pk);
  const auto& cols = pk.cols;
  hipStream_t st = s->stream;
  HIP_TRY(hipMemcpyAsync(pb.pos_mass.as<float4>() + at, pk.pm.data(), 16ull * n, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(pb.vel_vol.as<float4>() + at, pk.vv.data(), 16ull * n, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(pb.cls.as<uint32_t>() + at, pk.cls.data(), 4ull * n, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(pb.ins_id.as<uint32_t>() + at, pk.ins.data(), 4ull * n, hipMemcpyHostToDevice, st));
  for (int q = 0; q < 3; ++q)
    HIP_TRY(hipMemcpyAsync(pb.P.as<float>() + (size_t)q * s->cap + at, &cols[(size_t)(18 + q) * n], 4ull * n,
                           hipMemcpyHostToDevice, st));
  for (int q = 0; q < 9; ++q) {
    HIP_TRY(hipMemcpyAsync(pb.C.as<float>() + (size_t)q * s->cap + at, &cols[(size_t)q * n], 4ull * n,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(pb.F.as<float>() + (size_t)q * s->cap + at, &cols[(size_t)(9 + q) * n], 4ull * n,
                           hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return TMPM_OK;
}

// One-shot forces of scene k's particles, at their insertion indices.
tmpm_status upload_scene_forces(tmpm_session* s, uint32_t k) {
  SceneState& scene = s->scenes[k];
  const HostWorld& w = scene.world;
  if (w.size() == 0) return TMPM_OK;
  auto& fz = s->stage.forces;
  fz.resize(w.size());
  scene.force_pending = false;
  std::atomic<bool> bad{false}, pending{false};
  parallel_for(w.size(), [&](size_t b, size_t e) {
    bool any = false;
    for (size_t i = b; i < e; ++i) {
      float fx, fy, fzv;
      if (!finite_f32(w.force[3 * i], fx) || !finite_f32(w.force[3 * i + 1], fy) ||
          !finite_f32(w.force[3 * i + 2], fzv)) {
        bad = true;
        return;
      }
      fz[i] = make_float4(fx, fy, fzv, 0.f);
      any = any || (w.enabled[i] && (fx != 0.f || fy != 0.f || fzv != 0.f));
    }
    if (any) pending = true;
  });
  if (bad) return fail(TMPM_INVALID_INPUT, "force not representable in f32");
  scene.force_pending = pending;
  HIP_TRY(hipMemcpy(s->force.as<float4>() + scene.first, fz.data(), 16ull * w.size(), hipMemcpyHostToDevice));
  return TMPM_OK;
}

// Grows the force array (indexed by insertion index) to `need`, keeping its contents.
tmpm_status ensure_force_capacity(tmpm_session* s, uint32_t need) {
  if (need <= s->force_cap && s->force.ptr) return TMPM_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  const uint32_t cap = std::max<uint32_t>(std::max<uint32_t>(need, s->force_cap + s->force_cap / 2), 1024);
  DevBuf grown;
  HIP_TRY(grown.alloc(16ull * cap));
  HIP_TRY(hipMemset(grown.ptr, 0, 16ull * cap));
  if (s->force.ptr && s->force_cap)
    HIP_TRY(hipMemcpy(grown.ptr, s->force.ptr, 16ull * s->force_cap, hipMemcpyDeviceToDevice));
  std::swap(s->force.ptr, grown.ptr);
  std::swap(s->force.bytes, grown.bytes);
  s->force_cap = cap;
  return TMPM_OK;
}

tmpm_status upload_disabled(tmpm_session* s) {
  if (s->disabled_host.empty()) return TMPM_OK;
  HIP_TRY(s->disabled_rec.alloc(16ull * s->disabled_host.size()));
  HIP_TRY(hipMemcpy(s->disabled_rec.ptr, s->disabled_host.data(), 16ull * s->disabled_host.size(),
                    hipMemcpyHostToDevice));
  return TMPM_OK;
}

tmpm_status upload_classes(tmpm_session* s) {
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(s->classes.alloc(sizeof(DevClass) * std::max<size_t>(1, s->host_classes.size())));
  if (!s->host_classes.empty())
    HIP_TRY(hipMemcpy(s->classes.ptr, s->host_classes.data(), sizeof(DevClass) * s->host_classes.size(),
                      hipMemcpyHostToDevice));
  return TMPM_OK;
}

// Upload all active scenes to the device (gpu.rs:1091-1600 `prepare`,
// resident).  Insertion indices run scene by scene; the device holds the
// enabled particles in insertion order, batches with room to admit more.
tmpm_status prepare_packed(tmpm_session* s, PackedParticles& pk);

// Uploads every stepping scene.  A batch keeps the staging memory for later
// admissions; a plain session frees it.
tmpm_status prepare(tmpm_session* s) {
  s->stage.clear();
  tmpm_status st = prepare_packed(s, s->stage);
  if (!s->batch) s->stage.reset();
  return st;
}

tmpm_status prepare_packed(tmpm_session* s, PackedParticles& pk) {
  s->prepared = false;
  s->topo_valid = false;
  if (s->scenes.size() > 0xffffu) return fail(TMPM_CAPACITY, "too many scenes");
  uint64_t total = 0;
  for (auto& scene : s->scenes) {
    scene.first = (uint32_t)total;
    scene.count = scene.stepping() ? (uint32_t)std::min<size_t>(scene.world.size(), UINT32_MAX) : 0;
    scene.device_newer = false;
    total += scene.count;
    if (!scene.stepping()) continue;
    scene.max_mass = 0.0;
    for (size_t i = 0; i < scene.count; ++i)
      if (scene.world.enabled[i]) scene.max_mass = std::max(scene.max_mass, scene.world.mass[i]);
    tmpm_status st = build_scene(scene);
    if (st != TMPM_OK) return st;
  }
  if (total >= (1ull << 31)) return fail(TMPM_CAPACITY, "too many particles");
  s->id_space = (uint32_t)total;
  s->host_classes.clear();
  s->class_ids.clear();
  for (uint32_t k = 0; k < s->scenes.size(); ++k) {
    if (!s->scenes[k].stepping()) continue;
    tmpm_status st = pack_scene(s, k, pk);
    if (st != TMPM_OK) return st;
  }
  s->n = (uint32_t)pk.size();
  s->plastic = classes_plastic(s);
  s->n_enabled = s->n;
  s->synced_n = 0;
  s->disabled_host.assign(pk.disabled.begin(), pk.disabled.end());
  if (s->n == 0) {
    // gpu.rs:1358-1360: no live node, nothing to prepare; steps only count.
    return TMPM_OK;
  }
  s->cap = s->batch ? std::max<uint32_t>(s->n + s->n / 4, 4096) : s->n;
  for (int b = 0; b < 2; ++b) {
    tmpm_status st = alloc_particles(s->buf[b], s->cap);
    if (st != TMPM_OK) return st;
  }
  if (s->batch) {
    tmpm_status st = alloc_particles(s->synced, s->cap);
    if (st != TMPM_OK) return st;
  }
  s->cur = 0;
  tmpm_status st = upload_particles(s, pk, 0);
  if (st != TMPM_OK) return st;
  s->force_cap = 0;
  s->force.release();
  st = ensure_force_capacity(s, s->batch ? std::max<uint32_t>(s->id_space + s->id_space / 4, 4096) : s->id_space);
  if (st != TMPM_OK) return st;
  for (uint32_t k = 0; k < s->scenes.size(); ++k)
    if (s->scenes[k].stepping() && (st = upload_scene_forces(s, k)) != TMPM_OK) return st;
  if ((st = upload_disabled(s)) != TMPM_OK) return st;
  if ((st = upload_classes(s)) != TMPM_OK) return st;
  if ((st = upload_rigid(s)) != TMPM_OK) return st;
  if ((st = upload_obstacles(s)) != TMPM_OK) return st;
  if ((st = upload_scene(s)) != TMPM_OK) return st;
  HIP_TRY(hipMemset(s->scene_err.ptr, 0, 4ull * s->scene_cap));
  // topology capacity from the initial block counts, with headroom
  uint64_t pblocks = 0, gblocks = 0;
  count_blocks(pk.pos32, pk.region, pk.inv_h, pblocks, gblocks);
  s->last_pblocks = (uint32_t)pblocks;
  s->last_gblocks = (uint32_t)gblocks;
  // CPIC splits blocks that straddle a triangle by side: more headroom
  const uint64_t room = s->cpic ? 8 : 4;
  uint64_t pcap = std::max<uint64_t>(1024, room * pblocks);
  uint64_t gcap = std::max<uint64_t>(4096, room * gblocks);
  pcap = std::min<uint64_t>(pcap, s->cap);
  pcap = std::max<uint64_t>(pcap, 1);
  if (gcap > (1ull << 26)) return fail(TMPM_CAPACITY, "grid exceeds the block capacity");
  st = alloc_topology(s, (uint32_t)pcap, (uint32_t)gcap);
  if (st != TMPM_OK) return st;
  s->substeps_since_reorder = 1u << 30;
  s->compacted_last = false;
  s->prepared = true;
  return TMPM_OK;
}

// Binning stages of one substep (k_topo_clear .. k_pblock_finalize).
hipError_t enqueue_binning(tmpm_session* s, const Parts& p) {
  const DevClass* c = s->classes.as<DevClass>();
  const DevScene* sc = s->scene_plain.as<DevScene>();
  Topo& t = s->topo;
  hipError_t e;
  if (s->table_dirty) {
    if ((e = timed(s, "k_topo_clear", [&] { launch_topo_clear(t, s->stream); })) != hipSuccess) return e;
    s->table_dirty = false;
  }
  if ((uint64_t)t.bin_groups * t.bin_span < p.n) return hipErrorInvalidValue;  // lists sized for fewer particles
  if (s->cpic) {
    const DevObstacle* ob = s->obstacles.as<DevObstacle>();
    if ((e = timed(s, "k_cpic_mask", [&] { launch_cpic_mask(p, t, c, sc, ob, s->stream); })) != hipSuccess) return e;
  }
  if ((e = timed(s, "k_bin_insert", [&] { launch_bin_insert(p, t, c, sc, s->cpic, s->stream); })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_blocks", [&] { launch_bin_blocks(p, t, c, sc, s->cpic, s->stream); })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_offsets", [&] { launch_bin_offsets(p, t, c, sc, s->stream); })) != hipSuccess) return e;
  return hipSuccess;
}

// Everything one substep's launches depend on: two substeps with equal plans
// (compared bytewise, padding zeroed) launch the same kernels with the same
// arguments, so a graph captured for one replays the other.
struct SubstepPlan {
  Parts p, q;  // q: the buffer G2P moves the state into (zero unless reorder)
  Topo t;
  const DevClass* c;
  const DevScene* sc;       // this substep's scene table (with or without forces)
  const DevScene* sc_plain;
  DevObstacle* ob;
  DevRigid* rigid;
  const DevMapping* maps;
  const uint32_t* body_scene;
  uint32_t obstacle_total, body_total, p2g_blocks, p2g_threads, g2p_threads;
  int32_t react, fixed64, shapes, plastic, cpic, copies;
  double* audit;  // FP64 audit record set of this substep, or null
  uint8_t reorder, clear_table, rigid_sync, rigid_apply;
};

// Default launch shapes, from the particle-block count of the last
// synchronized substep and the mean particles per block (every shape computes
// the same result, reactions included: their particle-side sums are
// integers).  P2G and G2P run one workgroup per block, so with 64 threads a
// launch has one wave per block and each wave walks its block's particles
// serially.  Wider workgroups split a block when there are too few blocks to
// fill the chip (below kFillBlocks) or many particles per block:
//   G2P: the wider of 64 / 128 (below kFillBlocks) / 256 (below half of it)
//     and 64 / 128 (from 192 particles per block) / 256 (from 288).
//     Measured: 1M particles in 10,976 blocks of 96: 64 is fastest; 5.3M in
//     21,187 blocks of 252: 390 us at 128, 400 at 64 and 256; serving
//     batches of 3,123 blocks of 313 and 8,916 of 345: 112 -> 81 and
//     326 -> 284 us from 64 to 256 threads.
//   P2G: 128 threads below kFillBlocks or from 192 particles per block, else
//     64 (1M particles: 76 vs 86 us at 128; the batches above: 1-4% faster
//     at 128); one tile copy.  The reorder deals each block's particles
//     round-robin by cell, so a wavefront's lanes rarely share a cell, and a
//     second copy's LDS costs more workgroups per CU than the same-address
//     atomics it separates (1M particles: 61 vs 66 us).
constexpr uint32_t kFillBlocks = 8192;  // 32 waves per CU of MI355X's 256

uint32_t particles_per_block(const tmpm_session* s) {
  return s->last_pblocks ? (uint32_t)(s->n / s->last_pblocks) : 0u;
}

// Autotuner cache.  Tuned launch shapes are kept per workload bucket: device,
// particle count and particles per block (each rounded down to a power of
// two), and the features that select kernel variants (reactions, plastic
// materials, obstacle shapes, P2G mode, batch).  A session in a bucket that
// was tuned before takes the cached shape without trials, and sessions whose
// shape is not fixed (batches under admission and retirement among them)
// launch with the cached shape of the bucket they are in at each substep.
// The cache is process-wide; with TMPM_TUNE_CACHE naming a file it is also
// loaded from and appended to that file, one entry per line.
struct TuneKey {
  int device, n_log2, ppb_log2, features;
  bool operator<(const TuneKey& o) const {
    return std::tie(device, n_log2, ppb_log2, features) < std::tie(o.device, o.n_log2, o.ppb_log2, o.features);
  }
};
struct TuneShape {
  uint32_t p2g_threads = 64, g2p_threads = 64, reorder_interval = 16;
  int p2g_copies = 2, graphs = 0;
};

struct TuneCache {
  std::mutex mu;
  std::map<TuneKey, TuneShape> entries;
  bool loaded = false;
  std::string path;

  void load_locked() {
    if (loaded) return;
    loaded = true;
    const char* f = std::getenv("TMPM_TUNE_CACHE");
    if (!f || !*f) return;
    path = f;
    FILE* fp = std::fopen(f, "r");
    if (!fp) return;
    TuneKey k;
    TuneShape v;
    while (std::fscanf(fp, "%d %d %d %d %u %d %u %u %d", &k.device, &k.n_log2, &k.ppb_log2, &k.features,
                       &v.p2g_threads, &v.p2g_copies, &v.g2p_threads, &v.reorder_interval, &v.graphs) == 9) {
      const bool ok = (v.p2g_threads == 64 || v.p2g_threads == 128) && (v.p2g_copies == 1 || v.p2g_copies == 2) &&
                      (v.g2p_threads == 64 || v.g2p_threads == 128 || v.g2p_threads == 256) &&
                      v.reorder_interval >= 1 && v.reorder_interval <= 1000000;
      if (ok) entries[k] = v;
    }
    std::fclose(fp);
  }
  bool find(const TuneKey& k, TuneShape& out) {
    std::lock_guard<std::mutex> g(mu);
    load_locked();
    auto it = entries.find(k);
    if (it == entries.end()) return false;
    out = it->second;
    return true;
  }
  void put(const TuneKey& k, const TuneShape& v) {
    std::lock_guard<std::mutex> g(mu);
    load_locked();
    entries[k] = v;
    if (path.empty()) return;
    if (FILE* fp = std::fopen(path.c_str(), "a")) {
      std::fprintf(fp, "%d %d %d %d %u %d %u %u %d\n", k.device, k.n_log2, k.ppb_log2, k.features, v.p2g_threads,
                   v.p2g_copies, v.g2p_threads, v.reorder_interval, v.graphs);
      std::fclose(fp);
    }
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu);
    entries.clear();
  }
  size_t size() {
    std::lock_guard<std::mutex> g(mu);
    load_locked();
    return entries.size();
  }
};

TuneCache& tune_cache() {
  static TuneCache c;
  return c;
}

int floor_log2(uint64_t v) { return v ? 63 - __builtin_clzll(v) : -1; }

TuneKey tune_key(const tmpm_session* s) {
  TuneKey k;
  k.device = s->ctx ? s->ctx->device : 0;
  k.n_log2 = floor_log2(s->n);
  k.ppb_log2 = floor_log2(particles_per_block(s));
  k.features = (s->react_active() && s->obstacle_total > 0 ? 1 : 0) | (s->plastic ? 2 : 0) | (s->shapes ? 4 : 0) |
               (s->fixed64 ? 8 : 0) | (s->batch ? 16 : 0) | (s->cpic ? 32 : 0);
  return k;
}

// The cached shape of the session's current bucket, if it was tuned.
bool cached_shape(const tmpm_session* s, TuneShape& out) { return tune_cache().find(tune_key(s), out); }

uint32_t effective_g2p_threads(const tmpm_session* s) {
  if (s->g2p_shape_fixed) return s->threads_pblock;
  TuneShape c;
  if (cached_shape(s, c)) return c.g2p_threads;
  const uint32_t b = s->last_pblocks, ppb = particles_per_block(s);
  const uint32_t by_blocks = b >= kFillBlocks ? 64u : b >= kFillBlocks / 2 ? 128u : 256u;
  const uint32_t by_density = ppb >= 288 ? 256u : ppb >= 192 ? 128u : 64u;
  return std::max(by_blocks, by_density);
}

void effective_p2g_shape(const tmpm_session* s, uint32_t& threads, int& copies) {
  if (s->p2g_shape_fixed) {
    threads = s->p2g_threads;
    copies = s->p2g_copies;
    return;
  }
  TuneShape c;
  if (cached_shape(s, c)) {
    threads = c.p2g_threads;
    copies = c.p2g_copies;
    return;
  }
  const uint32_t b = s->last_pblocks;
  threads = b < kFillBlocks || particles_per_block(s) >= 192 ? 128u : 64u;
  copies = 1;
}

SubstepPlan plan_substep(tmpm_session* s, bool apply_force) {
  SubstepPlan pl;
  std::memset(&pl, 0, sizeof pl);
  pl.p = s->parts(s->cur);
  pl.reorder = s->substeps_since_reorder >= s->reorder_interval;
  if (pl.reorder) pl.q = s->parts(1 - s->cur);
  pl.t = s->topo;
  pl.c = s->classes.as<DevClass>();
  pl.sc = (apply_force ? s->scene_force : s->scene_plain).as<DevScene>();
  pl.sc_plain = s->scene_plain.as<DevScene>();
  pl.ob = s->obstacles.as<DevObstacle>();
  pl.rigid = s->rigid.as<DevRigid>();
  pl.maps = s->rigid_maps.as<DevMapping>();
  pl.body_scene = s->body_scene.as<uint32_t>();
  pl.obstacle_total = s->obstacle_total;
  pl.body_total = s->body_total;
  pl.p2g_blocks = s->p2g_blocks;
  effective_p2g_shape(s, pl.p2g_threads, pl.copies);
  pl.g2p_threads = effective_g2p_threads(s);
  pl.react = s->react_active() && s->obstacle_total > 0 ? 1 : 0;
  pl.fixed64 = s->fixed64 ? 1 : 0;
  pl.shapes = s->shapes;
  pl.plastic = s->plastic ? 1 : 0;
  pl.cpic = s->cpic;
  pl.clear_table = s->table_dirty;
  pl.rigid_sync = s->body_total > 0 && s->obstacle_total > 0;
  pl.rigid_apply = pl.react && s->body_total > 0;
  pl.audit = s->audit_on ? s->audit_next : nullptr;
  return pl;
}

// The launches of one substep, in stream order.  Changes no host state, so
// it can be recorded into a graph.
hipError_t launch_plan(tmpm_session* s, const SubstepPlan& pl) {
  const Topo& t = pl.t;
  const Parts& p = pl.p;
  hipStream_t st = s->stream;
  hipError_t e;
  if (pl.rigid_sync && (e = timed(s, "k_rigid_sync", [&] {
                          launch_rigid_sync(pl.rigid, pl.maps, pl.ob, pl.obstacle_total, st);
                        })) != hipSuccess)
    return e;
  if (pl.clear_table && (e = timed(s, "k_topo_clear", [&] { launch_topo_clear(t, st); })) != hipSuccess) return e;
  if (pl.cpic &&
      (e = timed(s, "k_cpic_mask", [&] { launch_cpic_mask(p, t, pl.c, pl.sc_plain, pl.ob, st); })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_insert", [&] { launch_bin_insert(p, t, pl.c, pl.sc_plain, pl.cpic, st); })) !=
      hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_blocks", [&] { launch_bin_blocks(p, t, pl.c, pl.sc_plain, pl.cpic, st); })) !=
      hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_offsets", [&] { launch_bin_offsets(p, t, pl.c, pl.sc_plain, st); })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_bin_scatter", [&] { launch_bin_scatter(p, t, st); })) != hipSuccess) return e;
  if ((e = timed(s, pl.fixed64 ? "k_p2g_fixed64" : "k_p2g", [&] {
         launch_p2g(p, t, pl.c, pl.sc, pl.fixed64, pl.p2g_blocks, pl.p2g_threads, pl.copies, pl.plastic, st);
       })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_grid_update",
                 [&] { launch_grid_update(t, pl.sc, pl.ob, pl.fixed64, pl.react, pl.shapes, st); })) != hipSuccess)
    return e;
  if ((e = timed(s, "k_g2p", [&] {
         // on a reorder substep G2P also moves the state into block order in q
         launch_g2p(p, pl.q, t, pl.c, pl.sc, pl.ob, pl.p2g_blocks, pl.g2p_threads, pl.react, pl.shapes, pl.plastic,
                    st);
       })) != hipSuccess)
    return e;
  if (pl.react && (e = timed(s, "k_react_finalize", [&] { launch_react_finalize(t, pl.ob, st); })) != hipSuccess)
    return e;
  if (pl.rigid_apply && (e = timed(s, "k_rigid_apply", [&] {
                           launch_rigid_apply(t, pl.sc, pl.rigid, pl.maps, pl.ob, pl.body_scene, pl.body_total, st);
                         })) != hipSuccess)
    return e;
  // the audit reads the state G2P wrote: the reorder target on a reorder substep
  if (pl.audit && (e = timed(s, "k_audit", [&] { launch_audit(pl.reorder ? pl.q : p, pl.c, pl.audit, st); })) !=
                      hipSuccess)
    return e;
  return hipSuccess;
}

// Replays the graph captured for an equal plan, or captures one.  At most
// kGraphCache graphs are kept (least recently used evicted): a steady run
// cycles through a few plans (buffer parity, reorder or not).
constexpr size_t kGraphCache = 8;

hipError_t launch_graph(tmpm_session* s, const SubstepPlan& pl) {
  for (auto& g : s->graphs)
    if (std::memcmp(g.plan.data(), &pl, sizeof pl) == 0) {
      g.last_use = ++s->graph_clock;
      return hipGraphLaunch(g.exec, s->stream);
    }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t le = launch_plan(s, pl);
  e = hipStreamEndCapture(s->stream, &graph);
  if (le != hipSuccess || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return le != hipSuccess ? le : e;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return e;
  if (s->graphs.size() >= kGraphCache) {
    auto lru = std::min_element(s->graphs.begin(), s->graphs.end(),
                                [](const CachedGraph& a, const CachedGraph& b) { return a.last_use < b.last_use; });
    (void)hipGraphExecDestroy(lru->exec);
    s->graphs.erase(lru);
  }
  CachedGraph g;
  g.plan.resize(sizeof pl);
  std::memcpy(g.plan.data(), &pl, sizeof pl);
  g.exec = exec;
  g.last_use = ++s->graph_clock;
  s->graphs.push_back(std::move(g));
  return hipGraphLaunch(exec, s->stream);
}

void destroy_graphs(tmpm_session* s) {
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

// One substep: its launches (eager, or a replayed graph when graphs are on
// and per-kernel timing is off), then the host-side bookkeeping.
hipError_t enqueue_substep(tmpm_session* s, bool apply_force) {
  const SubstepPlan pl = plan_substep(s, apply_force);
  if ((uint64_t)pl.t.bin_groups * pl.t.bin_span < pl.p.n) return hipErrorInvalidValue;  // lists sized for fewer particles
  const hipError_t e = s->use_graphs && !s->timing && !pl.audit ? launch_graph(s, pl) : launch_plan(s, pl);
  if (e != hipSuccess) return e;
  if (pl.clear_table) s->table_dirty = false;
  if (pl.reorder) {
    s->compacted_last = true;
    s->cur = 1 - s->cur;
    s->substeps_since_reorder = 0;
  } else {
    s->compacted_last = false;
  }
  s->substeps_since_reorder++;
  s->topo_valid = false;
  return hipSuccess;
}

// Room for `steps` more audit record sets on the device, zeroed on the
// session stream ahead of the substeps that fill them.  Growing with
// substeps in flight waits for them and keeps their records.
tmpm_status reserve_audit(tmpm_session* s, uint32_t steps) {
  const uint32_t stride = std::max<uint32_t>(s->scene_cap, 1);
  if (s->audit_used == 0) s->audit_stride = stride;
  if (s->audit_stride != stride) return fail(TMPM_READBACK, "audit layout changed with substeps pending");
  const size_t set = (size_t)stride * kAuditFields * sizeof(double);
  const uint64_t need = s->audit_used + steps;
  if (need > s->audit_cap) {
    const uint64_t cap = std::max<uint64_t>(need, std::max<uint64_t>(64, s->audit_cap * 2));
    DevBuf grown;
    HIP_TRY(grown.alloc(cap * set));
    if (s->audit_used) {
      HIP_TRY(hipStreamSynchronize(s->stream));
      HIP_TRY(hipMemcpy(grown.ptr, s->audit_dev.ptr, s->audit_used * set, hipMemcpyDeviceToDevice));
    }
    std::swap(s->audit_dev.ptr, grown.ptr);
    std::swap(s->audit_dev.bytes, grown.bytes);
    s->audit_cap = cap;
  }
  HIP_TRY(hipMemsetAsync(static_cast<char*>(s->audit_dev.ptr) + s->audit_used * set, 0, steps * set, s->stream));
  return TMPM_OK;
}

// After the device finished: the pending substeps' audit records to the host.
tmpm_status collect_audit(tmpm_session* s) {
  const size_t doubles = (size_t)s->audit_used * s->audit_stride * kAuditFields;
  s->audit_host.resize(doubles);
  if (doubles) HIP_TRY(hipMemcpy(s->audit_host.data(), s->audit_dev.ptr, doubles * sizeof(double), hipMemcpyDeviceToHost));
  s->audit_host_steps = s->audit_used;
  s->audit_host_stride = s->audit_stride;
  s->audit_used = 0;
  return TMPM_OK;
}

tmpm_status status_from_error_bits(uint32_t bits) {
  if (bits & (ERR_TOPOLOGY | ERR_FIXED_RANGE))
    return fail(TMPM_CAPACITY, (bits & ERR_TOPOLOGY) ? "sparse grid capacity exhausted"
                                                    : "fixed-point P2G contribution out of range");
  if (bits & ERR_RESIDENT)
    return fail(TMPM_RESIDENT_UNAVAILABLE,
                "a particle exceeded the CFL bound, left the grid or became non-finite; world state is unchanged");
  if (bits & ERR_INTERNAL) return fail(TMPM_READBACK, "internal binning inconsistency");
  return TMPM_OK;
}

// Copies device state into new worlds (gpu.rs:2295-2329 apply_gpu_transfers):
// out[k] is scene k's updated world and st[k] TMPM_NON_FINITE_STATE if one of
// its particles is not finite.  Scenes that are not stepping are left as is.
tmpm_status read_back(tmpm_session* s, std::vector<HostWorld>& out, std::vector<tmpm_status>& st) {
  const size_t ns = s->scenes.size();
  out.assign(ns, HostWorld());
  st.assign(ns, TMPM_OK);
  for (size_t k = 0; k < ns; ++k) {
    if (!s->scenes[k].stepping()) continue;
    HostWorld& w = out[k];
    w = s->scenes[k].world;
    // forces are one-shot; fixed particles end with zero v and C (world.rs:419-430)
    for (size_t i = 0; i < w.size(); ++i) {
      w.force[3 * i] = w.force[3 * i + 1] = w.force[3 * i + 2] = 0.0;
      if (w.fixed[i] && w.enabled[i]) {
        for (int q = 0; q < 3; ++q) w.vel[3 * i + q] = 0.0;
        for (int q = 0; q < 9; ++q) w.C[9 * i + q] = 0.0;
      }
    }
  }
  if (!s->prepared) return TMPM_OK;
  size_t n = s->n;
  const ParticleBuffers& b = s->buf[s->cur];
  std::vector<float4> pm(n), vv(n);
  std::vector<float> C(9 * n), F(9 * n);
  std::vector<uint32_t> ins(n);
  HIP_TRY(hipMemcpy(pm.data(), b.pos_mass.ptr, 16 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(vv.data(), b.vel_vol.ptr, 16 * n, hipMemcpyDeviceToHost));
  for (int q = 0; q < 9; ++q) {
    HIP_TRY(hipMemcpy(C.data() + (size_t)q * n, b.C.as<float>() + (size_t)q * s->cap, 4 * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(F.data() + (size_t)q * n, b.F.as<float>() + (size_t)q * s->cap, 4 * n, hipMemcpyDeviceToHost));
  }
  HIP_TRY(hipMemcpy(ins.data(), b.ins_id.ptr, 4 * n, hipMemcpyDeviceToHost));
  std::vector<float> P(s->plastic ? 3 * n : 0);
  for (int q = 0; q < 3 && s->plastic; ++q)
    HIP_TRY(hipMemcpy(P.data() + (size_t)q * n, b.P.as<float>() + (size_t)q * s->cap, 4 * n, hipMemcpyDeviceToHost));
  // insertion index -> scene: scenes occupy consecutive ranges
  std::vector<std::pair<uint32_t, uint32_t>> ranges;  // (first, scene)
  for (uint32_t k = 0; k < ns; ++k)
    if (s->scenes[k].count) ranges.emplace_back(s->scenes[k].first, k);
  for (size_t slot = 0; slot < n; ++slot) {
    const uint32_t g = ins[slot];
    if (g == kInvalid) continue;  // retired slot
    auto it = std::upper_bound(ranges.begin(), ranges.end(), std::make_pair(g, UINT32_MAX));
    if (it == ranges.begin()) return fail(TMPM_READBACK, "corrupt particle order");
    const uint32_t k = (it - 1)->second;
    const size_t i = g - s->scenes[k].first;
    HostWorld& w = out[k];
    if (i >= w.size() || i >= s->scenes[k].count) return fail(TMPM_READBACK, "corrupt particle order");
    if (!w.enabled[i] || w.fixed[i]) continue;
    const float4 p = pm[slot], v = vv[slot];
    double xs[3] = {p.x, p.y, p.z}, vs[3] = {v.x, v.y, v.z};
    bool finite = true;
    for (int q = 0; q < 3; ++q) {
      w.pos[3 * i + q] = xs[q];
      w.vel[3 * i + q] = vs[q];
      finite = finite && std::isfinite(xs[q]) && std::isfinite(vs[q]);
    }
    for (int q = 0; q < 9; ++q) {
      w.C[9 * i + q] = C[q * n + slot];
      w.F[9 * i + q] = F[q * n + slot];
      finite = finite && std::isfinite(w.C[9 * i + q]) && std::isfinite(w.F[9 * i + q]);
    }
    if (s->plastic && is_plastic(w.material[i]))
      for (int q = 0; q < 3; ++q) w.plastic[3 * i + q] = P[q * n + slot];
    if (!finite) st[k] = TMPM_NON_FINITE_STATE;
  }
  return TMPM_OK;
}

// Batches: brings scene k's host world up to its last synchronize by
// gathering its particles from the synchronized copy on the device.
tmpm_status materialize(tmpm_session* s, uint32_t k) {
  SceneState& scene = s->scenes[k];
  if (!scene.device_newer) return TMPM_OK;
  if (!s->prepared || s->synced_n == 0 || scene.count == 0) {
    scene.device_newer = false;
    return TMPM_OK;
  }
  const size_t bytes = 4ull * kGatherFloats * scene.count;
  if (s->staging.bytes < bytes) HIP_TRY(s->staging.alloc(bytes));
  Parts sp = s->parts(0);
  sp.pos_mass = s->synced.pos_mass.as<float4>();
  sp.vel_vol = s->synced.vel_vol.as<float4>();
  sp.C = s->synced.C.as<float>();
  sp.F = s->synced.F.as<float>();
  sp.P = s->plastic ? s->synced.P.as<float>() : nullptr;
  sp.cls = s->synced.cls.as<uint32_t>();
  sp.ins_id = s->synced.ins_id.as<uint32_t>();
  sp.n = s->synced_n;
  launch_gather_scene(sp, scene.first, scene.count, s->staging.as<float>(), s->stream);
  HIP_TRY(hipGetLastError());
  std::vector<float> rec((size_t)kGatherFloats * scene.count);
  HIP_TRY(hipMemcpyAsync(rec.data(), s->staging.ptr, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  HostWorld w = scene.world;
  for (size_t i = 0; i < w.size(); ++i) {
    w.force[3 * i] = w.force[3 * i + 1] = w.force[3 * i + 2] = 0.0;
    if (!w.enabled[i]) continue;
    if (w.fixed[i]) {
      for (int q = 0; q < 3; ++q) w.vel[3 * i + q] = 0.0;
      for (int q = 0; q < 9; ++q) w.C[9 * i + q] = 0.0;
      continue;
    }
    const float* r = &rec[(size_t)kGatherFloats * i];
    for (int q = 0; q < 3; ++q) {
      w.pos[3 * i + q] = r[q];
      w.vel[3 * i + q] = r[4 + q];
    }
    for (int q = 0; q < 9; ++q) {
      w.C[9 * i + q] = r[8 + q];
      w.F[9 * i + q] = r[17 + q];
    }
    if (s->plastic && is_plastic(w.material[i]))
      for (int q = 0; q < 3; ++q) w.plastic[3 * i + q] = r[26 + q];
  }
  const uint64_t substeps = scene.world.substeps;
  scene.set_world(w);
  scene.world.substeps = substeps;
  scene.device_newer = false;
  return TMPM_OK;
}

// Grows the topology when the last substep used more than half of it.
tmpm_status maybe_grow_topology(tmpm_session* s, const uint32_t* counters) {
  uint32_t g = counters[CNT_GRID_BLOCKS], pbk = counters[CNT_PARTICLE_BLOCKS];
  if (2ull * g <= s->topo.grid_cap && 2ull * pbk <= s->topo.pblock_cap) return TMPM_OK;
  uint64_t gcap = std::max<uint64_t>(s->topo.grid_cap, 4ull * g);
  uint64_t pcap = std::min<uint64_t>(std::max<uint64_t>(s->topo.pblock_cap, 4ull * pbk), std::max<uint32_t>(s->cap, 1));
  return alloc_topology(s, (uint32_t)pcap, (uint32_t)std::min<uint64_t>(gcap, 1ull << 26));
}

tmpm_status copy_particles_out(const HostWorld& w, tmpm_particles* out) {
  size_t n = std::min<size_t>(out->count, w.size());
  auto put = [n](double* dst, const DVec& src, int k) {
    if (dst) std::memcpy(dst, src.data(), sizeof(double) * k * n);
  };
  put(out->position, w.pos, 3);
  put(out->velocity, w.vel, 3);
  put(out->deformation, w.F, 9);
  put(out->affine, w.C, 9);
  put(out->force, w.force, 3);
  put(out->mass, w.mass, 1);
  put(out->rest_volume, w.rest_volume, 1);
  put(out->radius, w.radius, 1);
  put(out->damping, w.damping, 1);
  put(out->plastic, w.plastic, 3);
  for (size_t i = 0; i < n; ++i) {
    if (out->group_id) out->group_id[i] = w.group_id[i];
    if (out->transfer_color) out->transfer_color[i] = w.color[i];
    if (out->enabled) out->enabled[i] = w.enabled[i];
    if (out->fixed) out->fixed[i] = w.fixed[i];
    if (out->chunk_id) out->chunk_id[i] = w.chunk[i];
    if (out->material_index) out->material_index[i] = 0;
  }
  return TMPM_OK;
}

tmpm_status run_diagnostic_binning(tmpm_session* s, std::vector<uint32_t>& counters) {
  counters.assign(CNT_COUNT, 0);
  if (!s->prepared) return TMPM_OK;
  Parts p = s->parts(s->cur);
  bool was = s->timing;
  s->timing = false;
  HIP_TRY(enqueue_binning(s, p));
  launch_bin_scatter(p, s->topo, s->stream);  // fills the per-block tile slots
  s->table_dirty = true;
  launch_live_nodes(p, s->topo, s->classes.as<DevClass>(), s->scene_plain.as<DevScene>(), s->stream);
  s->timing = was;
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(counters.data(), s->counters.ptr, 4 * CNT_COUNT, hipMemcpyDeviceToHost));
  s->topo_valid = true;
  return TMPM_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// Exceptions never cross the C ABI: every status-returning entry point
// catches them here.  A host allocation failure is Capacity; anything else
// is an internal Readback error.  The session (if any) is invalidated, since
// the failed call may have stopped halfway; its last synchronized world stays
// readable.
tmpm_status exception_status() {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(TMPM_CAPACITY, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(TMPM_READBACK, std::string("internal error: ") + e.what());
  } catch (...) {
    return fail(TMPM_READBACK, "internal error");
  }
}
tmpm_status api_exception(tmpm_session* s) {
  if (s) s->valid = false;
  return exception_status();
}
// read-only entry points leave the session as it was
tmpm_status api_exception(const tmpm_session*) { return exception_status(); }

extern "C" {

const char* tmpm_last_error(void) { return g_last_error.c_str(); }
const char* tmpm_version(void) { return "tessera-mpm-hip 0.1.0 (gfx950)"; }

}  // extern "C"

namespace {

// Scratch of tmpm_sample_closed_mesh_volume_noalloc, the entry point that
// allocates nothing: a status word, the mesh's vertices and triangles and one
// inside flag per lattice candidate.  Each device holds its own copy for the
// life of the process; in host code only its per-device address (from
// hipGetSymbolAddress, set up when a context is created on the device) may be
// passed to kernels or copies.
constexpr uint64_t kScratchBytes = 32ull << 20;
constexpr uint64_t kNoAllocMaxTriangles = 1ull << 16;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocScratch {
  std::mutex mu;         // serializes the calls of this process on the device
  void* ptr = nullptr;   // the device's address of `scratch`
};

NoAllocScratch& noalloc_scratch(int device) {
  static NoAllocScratch table[256];
  return table[device & 255];
}

// Looks up this device's address of `scratch`; the device must be current.
tmpm_status setup_scratch(int device) {
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(TMPM_READBACK, std::string("cannot find the sampler's scratch memory: ") + hipGetErrorString(e));
  noalloc_scratch(device).ptr = ptr;
  return TMPM_OK;
}

// Argument, vertex and triangle checks of sampling.rs:54-93 shared by both
// samplers, all but the edge test: geometry, spacing and limits; finite
// vertices; triangles of three distinct in-range vertices and non-zero area.
tmpm_status mesh_check_triangles(tmpm_context* ctx, const double* vertices, uint64_t vertex_count,
                                 const uint32_t* triangles, uint64_t triangle_count, double spacing,
                                 uint64_t max_candidates, uint64_t max_samples, double* samples_out,
                                 uint64_t* count_out) {
  if (!ctx || !count_out) return fail(TMPM_INVALID_INPUT, "null argument");
  if (vertex_count == 0 || triangle_count == 0 || !vertices || !triangles || !samples_out ||
      !std::isfinite(spacing) || spacing <= 0.0 || max_candidates == 0 || max_samples == 0)
    return fail(TMPM_INVALID_INPUT, "invalid mesh, spacing or limits");
  for (uint64_t i = 0; i < 3 * vertex_count; ++i)
    if (!std::isfinite(vertices[i])) return fail(TMPM_INVALID_INPUT, "non-finite mesh vertex");
  for (uint64_t t = 0; t < triangle_count; ++t) {
    const uint32_t* tr = triangles + 3 * t;
    if (tr[0] >= vertex_count || tr[1] >= vertex_count || tr[2] >= vertex_count || tr[0] == tr[1] ||
        tr[1] == tr[2] || tr[2] == tr[0])
      return fail(TMPM_INVALID_INPUT, "invalid mesh triangle");
    const double* a = vertices + 3ull * tr[0];
    const double* b = vertices + 3ull * tr[1];
    const double* c = vertices + 3ull * tr[2];
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                          e1[0] * e2[1] - e1[1] * e2[0]};
    const double area2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
    if (!std::isfinite(area2) || area2 <= 1e-24) return fail(TMPM_INVALID_INPUT, "degenerate mesh triangle");
  }
  return TMPM_OK;
}

struct MeshLattice {
  double lower[3], upper[3];
  uint64_t counts[3];
  uint64_t candidates;
};

// sampling.rs:94-115: the AABB lattice, bounded by max_candidates.
tmpm_status mesh_lattice(const double* vertices, uint64_t vertex_count, double spacing, uint64_t max_candidates,
                         MeshLattice& lat) {
  for (int a = 0; a < 3; ++a) lat.lower[a] = lat.upper[a] = vertices[a];
  for (uint64_t i = 1; i < vertex_count; ++i)
    for (int a = 0; a < 3; ++a) {
      lat.lower[a] = std::fmin(lat.lower[a], vertices[3 * i + a]);
      lat.upper[a] = std::fmax(lat.upper[a], vertices[3 * i + a]);
    }
  lat.candidates = 1;
  for (int a = 0; a < 3; ++a) {
    const double c = std::ceil((lat.upper[a] - lat.lower[a]) / spacing);
    if (!std::isfinite(c) || c < 1.0 || c > (double)max_candidates)
      return fail(TMPM_CAPACITY, "candidate lattice exceeds max_candidates");
    lat.counts[a] = (uint64_t)c;
    if (lat.candidates > max_candidates / lat.counts[a])
      return fail(TMPM_CAPACITY, "candidate lattice exceeds max_candidates");
    lat.candidates *= lat.counts[a];
  }
  return TMPM_OK;
}

// Accepted centres of candidates [c0, c0 + len) (flags[c - c0] != 0) appended
// to samples_out in lattice order (x, then y, then z), at most max_samples.
tmpm_status mesh_emit_samples(const MeshLattice& lat, double spacing, const uint8_t* flags, uint64_t c0,
                              uint64_t len, uint64_t max_samples, double* samples_out, uint64_t& n) {
  for (uint64_t c = c0; c < c0 + len; ++c) {
    if (!flags[c - c0]) continue;
    if (n == max_samples) return fail(TMPM_CAPACITY, "accepted samples exceed max_samples");
    const uint64_t x = c / (lat.counts[1] * lat.counts[2]), y = (c / lat.counts[2]) % lat.counts[1],
                   z = c % lat.counts[2];
    samples_out[3 * n] = lat.lower[0] + ((double)x + 0.5) * spacing;
    samples_out[3 * n + 1] = lat.lower[1] + ((double)y + 0.5) * spacing;
    samples_out[3 * n + 2] = lat.lower[2] + ((double)z + 0.5) * spacing;
    ++n;
  }
  return TMPM_OK;
}

}  // namespace

extern "C" {

tmpm_status tmpm_context_create(int32_t device_ordinal, tmpm_context** out) {
  try {
    if (!out) return fail(TMPM_INVALID_INPUT, "null output");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device_ordinal < 0 || device_ordinal >= count)
      return fail(TMPM_NO_DEVICE, "no HIP device with this ordinal");
    tmpm_context* c = new tmpm_context();
    c->device = device_ordinal;
    if (hipGetDeviceProperties(&c->props, device_ordinal) != hipSuccess) {
      delete c;
      return fail(TMPM_NO_DEVICE, "cannot query device");
    }
    if (std::strncmp(c->props.gcnArchName, "gfx950", 6) != 0) {
      std::string arch = c->props.gcnArchName;
      delete c;
      return fail(TMPM_NO_DEVICE, "device is " + arch + ", this library is built for gfx950");
    }
    {
      // once per device: the device address of the allocation-free sampler's scratch
      NoAllocScratch& ns = noalloc_scratch(device_ordinal);
      std::lock_guard<std::mutex> lock(ns.mu);
      if (!ns.ptr) {
        int prev = 0;
        const bool restore = hipGetDevice(&prev) == hipSuccess;
        tmpm_status st = hipSetDevice(device_ordinal) == hipSuccess ? setup_scratch(device_ordinal)
                                                                    : fail(TMPM_NO_DEVICE, "cannot select device");
        if (restore) (void)hipSetDevice(prev);
        if (st != TMPM_OK) {
          delete c;
          return st;
        }
      }
    }
    *out = c;
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

void tmpm_context_destroy(tmpm_context* ctx) { delete ctx; }

}  // extern "C"

namespace {

// Reactions, rigid bodies and the poses of body-driven obstacles, read back
// for the scenes whose substeps were committed.
tmpm_status read_coupling(tmpm_session* s, const std::vector<char>& commit) {
  if (!s->prepared || s->obstacle_total == 0) return TMPM_OK;
  if (s->react_active()) {
    std::vector<double> total(6ull * s->obstacle_total);
    HIP_TRY(hipMemcpy(total.data(), s->react_total.ptr, 8ull * total.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(s->react_total.ptr, 0, 8ull * total.size()));
    for (size_t k = 0; k < s->scenes.size(); ++k) {
      SceneState& scene = s->scenes[k];
      if (!commit[k]) continue;
      const size_t b = 6ull * scene.dev.obstacle_begin, n = 6ull * scene.dev.obstacle_count;
      scene.reactions.assign(total.begin() + b, total.begin() + b + n);
    }
  }
  if (s->body_total > 0) {
    std::vector<DevRigid> rb(s->body_total);
    std::vector<DevObstacle> ob(s->obstacle_total);
    HIP_TRY(hipMemcpy(rb.data(), s->rigid.ptr, sizeof(DevRigid) * rb.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ob.data(), s->obstacles.ptr, sizeof(DevObstacle) * ob.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < s->scenes.size(); ++k) {
      SceneState& scene = s->scenes[k];
      if (!commit[k] || scene.bodies.empty()) continue;
      for (size_t i = 0; i < scene.bodies.size(); ++i) {
        const DevRigid& d = rb[scene.dev.body_begin + i];
        tmpm_rigid_body& b = scene.bodies[i];
        for (int a = 0; a < 3; ++a) {
          b.position[a] = d.pos[a];
          b.linear_velocity[a] = d.lin[a];
          b.angular_velocity[a] = d.ang[a];
        }
        for (int a = 0; a < 4; ++a) b.orientation[a] = d.q[a];
      }
      for (size_t j = 0; j < scene.world.obstacles.size() && j < scene.maps.size(); ++j) {
        if (scene.maps[j].body == UINT32_MAX) continue;
        const DevObstacle& d = ob[scene.dev.obstacle_begin + j];
        tmpm_obstacle& o = scene.world.obstacles[j].o;
        for (int a = 0; a < 3; ++a) {
          o.center[a] = d.center[a];
          o.linear_velocity[a] = d.lin[a];
          o.angular_velocity[a] = d.ang[a];
        }
        for (int a = 0; a < 4; ++a) o.orientation[a] = d.q[a];
      }
    }
  }
  return TMPM_OK;
}

tmpm_status read_scene_reactions(const SceneState& scene, double* out, uint32_t cap, uint32_t* count_out) {
  const uint32_t n = (uint32_t)scene.world.obstacles.size();
  if (count_out) *count_out = n;
  if (cap < n || (n > 0 && !out)) return fail(TMPM_CAPACITY, "reaction output too small");
  for (uint32_t i = 0; i < 6 * n; ++i) out[i] = i < scene.reactions.size() ? scene.reactions[i] : 0.0;
  return TMPM_OK;
}

tmpm_status read_scene_bodies(const SceneState& scene, tmpm_rigid_body* out, uint32_t cap, uint32_t* count_out) {
  const uint32_t n = (uint32_t)scene.bodies.size();
  if (count_out) *count_out = n;
  if (cap < n || (n > 0 && !out)) return fail(TMPM_CAPACITY, "rigid body output too small");
  for (uint32_t i = 0; i < n; ++i) out[i] = scene.bodies[i];
  return TMPM_OK;
}

tmpm_status read_scene_obstacles(const SceneState& scene, tmpm_obstacle* out, uint32_t cap, uint32_t* count_out) {
  const uint32_t n = (uint32_t)scene.world.obstacles.size();
  if (count_out) *count_out = n;
  if (cap < n || (n > 0 && !out)) return fail(TMPM_CAPACITY, "obstacle output too small");
  for (uint32_t i = 0; i < n; ++i) {
    out[i] = scene.world.obstacles[i].o;
    out[i].planes = nullptr;
  }
  return TMPM_OK;
}

tmpm_status set_scene_rigid(tmpm_session* s, uint32_t k, const tmpm_rigid_body* bodies, uint32_t body_count,
                            const tmpm_rigid_mapping* maps, uint32_t map_count, const double gravity[3],
                            double max_speed, int32_t integrate) {
  SceneState& scene = s->scenes[k];
  if (!scene.stepping()) return fail(scene.status == TMPM_OK ? TMPM_INVALID_INPUT : scene.status, "scene is stopped");
  if ((body_count > 0 && (!bodies || !maps || !gravity)) ) return fail(TMPM_INVALID_INPUT, "null argument");
  if (body_count > 0 && map_count != scene.world.obstacles.size())
    return fail(TMPM_INVALID_INPUT, "one rigid mapping per obstacle is required");
  if (body_count > 0) {
    if (!std::isfinite(max_speed) || max_speed < 0.0) return fail(TMPM_INVALID_INPUT, "invalid speed cap");
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(gravity[a])) return fail(TMPM_INVALID_INPUT, "invalid gravity");
    for (uint32_t i = 0; i < body_count; ++i) {
      DevRigid d;
      if (!pack_rigid(bodies[i], d)) return fail(TMPM_INVALID_INPUT, "invalid rigid body");
    }
    for (uint32_t j = 0; j < map_count; ++j)
      if (maps[j].body != UINT32_MAX && maps[j].body >= body_count)
        return fail(TMPM_INVALID_INPUT, "rigid mapping names no body");
  }
  // rollback keeps only this scene's coupling (copying whole scenes would copy their particles)
  auto saved_bodies = scene.bodies;
  auto saved_maps = scene.maps;
  double saved_gravity[3] = {scene.rigid_gravity[0], scene.rigid_gravity[1], scene.rigid_gravity[2]};
  const double saved_speed = scene.rigid_max_speed;
  const bool saved_integrate = scene.rigid_integrate;
  scene.bodies.assign(bodies, bodies + body_count);
  scene.maps.assign(maps, maps + (body_count > 0 ? map_count : 0));
  if (body_count > 0) {
    for (int a = 0; a < 3; ++a) scene.rigid_gravity[a] = gravity[a];
    scene.rigid_max_speed = max_speed;
    scene.rigid_integrate = integrate != 0;
  }
  if (!s->prepared) return TMPM_OK;
  tmpm_status st = upload_rigid(s);
  if (st == TMPM_OK) st = upload_obstacles(s);
  if (st == TMPM_OK) st = upload_scene(s);
  if (st != TMPM_OK) {
    std::string msg = g_last_error;
    scene.bodies.swap(saved_bodies);
    scene.maps.swap(saved_maps);
    for (int a = 0; a < 3; ++a) scene.rigid_gravity[a] = saved_gravity[a];
    scene.rigid_max_speed = saved_speed;
    scene.rigid_integrate = saved_integrate;
    if (upload_rigid(s) != TMPM_OK || upload_obstacles(s) != TMPM_OK || upload_scene(s) != TMPM_OK) s->valid = false;
    g_last_error = msg;
  }
  return st;
}

tmpm_status check_world_desc(const tmpm_world_desc* wd) {
  if (!wd) return fail(TMPM_INVALID_INPUT, "null world");
  if (!params_valid(wd->params)) return fail(TMPM_INVALID_INPUT, "invalid MPM parameters");
  if (!validate_particles(wd->particles, wd->materials, wd->material_count))
    return fail(TMPM_INVALID_INPUT, "invalid MPM particle");
  for (uint32_t i = 0; i < wd->obstacle_count; ++i)
    if (!obstacle_valid(wd->obstacles[i])) return fail(TMPM_INVALID_INPUT, "invalid obstacle");
  return TMPM_OK;
}

// Loads a world description into `scene`; `w` may be an emptied world whose
// memory is reused.
void load_scene(SceneState& scene, const tmpm_world_desc* wd, double substep_dt, HostWorld w = HostWorld()) {
  w.params = wd->params;
  w.append(wd->particles, wd->materials, 0);
  // the world keeps its chunks (GpuMpmResidentSession takes the MpmWorld by value)
  uint64_t top = 0;
  if (wd->particles.chunk_id) {
    const size_t base = w.size() - wd->particles.count;
    for (uint64_t i = 0; i < wd->particles.count; ++i) {
      w.chunk[base + i] = wd->particles.chunk_id[i];
      top = std::max(top, wd->particles.chunk_id[i]);
    }
  }
  w.next_chunk_id = wd->next_chunk_id ? wd->next_chunk_id : (top == UINT64_MAX ? top : top + 1);
  w.set_obstacles(wd->obstacles, wd->obstacle_count);
  w.substeps = wd->substeps;
  scene = SceneState();
  scene.set_world(std::move(w));
  scene.substep_dt = substep_dt;
}

tmpm_status create_session(tmpm_context* ctx, const tmpm_world_desc* worlds, const double* dts, uint32_t count,
                           tmpm_p2g_mode mode, bool batch, tmpm_session** out) {
  if (!ctx || !out || (count > 0 && (!worlds || !dts))) return fail(TMPM_INVALID_INPUT, "null argument");
  *out = nullptr;
  if (mode != TMPM_P2G_FP32 && mode != TMPM_P2G_FIXED64) return fail(TMPM_INVALID_INPUT, "unknown P2G mode");
  if (count > 0xffffu) return fail(TMPM_CAPACITY, "too many scenes");
  for (uint32_t k = 0; k < count; ++k) {
    tmpm_status st = check_world_desc(&worlds[k]);
    if (st != TMPM_OK) return st;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(TMPM_NO_DEVICE, "cannot select device");
  tmpm_session* s = new tmpm_session();
  s->ctx = ctx;
  s->batch = batch;
  s->fixed64 = mode == TMPM_P2G_FIXED64;
  read_tuning_env(s);
  s->scenes.resize(count);
  for (uint32_t k = 0; k < count; ++k) load_scene(s->scenes[k], &worlds[k], dts[k]);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(TMPM_READBACK, "cannot create stream");
  }
  tmpm_status st = TMPM_OK;
  for (auto& scene : s->scenes)
    if (st == TMPM_OK) st = validate_resident_world(scene);
  if (st == TMPM_OK) st = prepare(s);
  if (st != TMPM_OK) {
    std::string msg = g_last_error;
    tmpm_session_destroy(s);
    g_last_error = msg;
    return st;
  }
  *out = s;
  return TMPM_OK;
}

// Common checks of an edit: the session (or the batch scene) must be usable.
tmpm_status check_editable(tmpm_session* s) {
  if (!s) return fail(TMPM_INVALID_INPUT, "null session");
  if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
  if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
  return TMPM_OK;
}

tmpm_status check_scene(const tmpm_session* s, uint32_t scene) {
  if (!s) return fail(TMPM_INVALID_INPUT, "null session");
  if (scene >= s->scenes.size() || !s->scenes[scene].active) return fail(TMPM_INVALID_INPUT, "no such scene");
  return TMPM_OK;
}

// Re-uploads every stepping scene after a host-side change; on failure the
// previous scene states are restored (edits are transactional).
tmpm_status reprepare(tmpm_session* s, std::vector<SceneState>& previous) {
  // prepare uploads host worlds: bring lazily synchronized scenes up to date first
  for (uint32_t k = 0; k < previous.size() && k < s->scenes.size(); ++k) {
    if (!previous[k].device_newer) continue;
    std::swap(s->scenes[k], previous[k]);
    tmpm_status st = materialize(s, k);
    std::swap(s->scenes[k], previous[k]);
    if (st != TMPM_OK) return st;
    if (s->scenes[k].active && s->scenes[k].first == previous[k].first) {
      s->scenes[k].world = previous[k].world;
      s->scenes[k].stable = -1.0;
      s->scenes[k].device_newer = false;
    }
  }
  tmpm_status st = prepare(s);
  if (st != TMPM_OK) {
    std::string msg = g_last_error;
    s->scenes.swap(previous);
    if (prepare(s) != TMPM_OK) s->valid = false;
    g_last_error = msg;
  }
  return st;
}

// Packed-particle staging buffers (page-locked) reused across admissions:
// freeing page-locked memory can wait for the device, so staged scenes
// return their buffers here instead.
struct PackPool {
  std::mutex mu;
  std::vector<std::unique_ptr<PackedParticles>> free;
  std::unique_ptr<PackedParticles> acquire() {
    std::lock_guard<std::mutex> g(mu);
    if (free.empty()) return std::make_unique<PackedParticles>();
    auto p = std::move(free.back());
    free.pop_back();
    return p;
  }
  void release(std::unique_ptr<PackedParticles> p) {
    if (!p) return;
    p->clear();
    std::lock_guard<std::mutex> g(mu);
    if (free.size() < 8) free.push_back(std::move(p));
  }
};

PackPool& pack_pool() {
  static PackPool* pool = new PackPool();  // never destroyed: no page-locked free at exit
  return *pool;
}

// Host half of an admission (no session, no device): validation, the f64
// copy of the world, f32 packing with scene index 0 and insertion base 0,
// and the block counts for the capacity check.  `spare`, when given, is an
// emptied world whose memory is reused.
tmpm_status stage_scene(const tmpm_world_desc* world, double substep_dt, HostWorld* spare, tmpm_staged_scene& out) {
  if (!world) return fail(TMPM_INVALID_INPUT, "null world");
  if (!params_valid(world->params)) return fail(TMPM_INVALID_INPUT, "invalid MPM parameters");
  for (uint32_t i = 0; i < world->obstacle_count; ++i)
    if (!obstacle_valid(world->obstacles[i])) return fail(TMPM_INVALID_INPUT, "invalid obstacle");
  const tmpm_particles& pp = world->particles;
  if (pp.count > 0 && (!pp.position || !pp.mass || !pp.rest_volume || !pp.radius || world->material_count == 0))
    return fail(TMPM_INVALID_INPUT, "invalid MPM particle");
  for (uint64_t i = 0; pp.material_index && i < pp.count; ++i)
    if (pp.material_index[i] >= world->material_count) return fail(TMPM_INVALID_INPUT, "invalid MPM particle");
  SceneState& scene = out.scene;
  if (spare) load_scene(scene, world, substep_dt, std::move(*spare));
  else load_scene(scene, world, substep_dt);
  {
    std::vector<char> bad(64, 0);
    parallel_for(scene.world.size(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i)
        if (!scene.world.particle_valid(i)) {
          bad[(b / 16384) % 64] = 1;
          break;
        }
    });
    for (char c : bad)
      if (c) return fail(TMPM_INVALID_INPUT, "invalid MPM particle");
  }
  tmpm_status st = validate_resident_world(scene);
  if (st != TMPM_OK) return st;
  if (scene.world.size() >= (1ull << 31)) return fail(TMPM_CAPACITY, "too many particles");
  scene.first = 0;
  scene.count = (uint32_t)scene.world.size();
  for (size_t i = 0; i < scene.count; ++i)
    if (scene.world.enabled[i]) scene.max_mass = std::max(scene.max_mass, scene.world.mass[i]);
  st = build_scene(scene);
  if (st != TMPM_OK) return st;
  out.pk = pack_pool().acquire();
  std::vector<DevClass> classes;
  st = pack_scene_host(scene, 0, *out.pk, classes);
  if (st != TMPM_OK) return st;
  // The scene's distinct classes, in the order interning would meet them
  // (runs of particles share a class: compare with the previous one first),
  // so that admission interns a few classes instead of every particle's.
  out.class_local.resize(out.pk->size());
  {
    std::map<ClassKey, uint32_t> seen;
    uint32_t last = UINT32_MAX;
    size_t e = 0;
    const HostWorld& w = scene.world;
    for (size_t i = 0; i < w.size(); ++i) {
      if (!w.enabled[i]) continue;
      const DevClass& c = classes[i];
      uint32_t id;
      if (last != UINT32_MAX && std::memcmp(&out.class_table[last], &c, sizeof(DevClass)) == 0) {
        id = last;
      } else {
        auto key = class_key(c);
        auto it = seen.find(key);
        if (it != seen.end()) {
          id = it->second;
        } else {
          id = (uint32_t)out.class_table.size();
          out.class_table.push_back(c);
          seen.emplace(key, id);
        }
      }
      out.class_local[e++] = id;
      last = id;
    }
  }
  count_blocks(out.pk->pos32, out.pk->region, out.pk->inv_h, out.pblocks, out.gblocks);
  pack_columns(*out.pk);
  return TMPM_OK;
}

// Device half of an admission into a batch with no pending substeps: the
// staged scene takes the first free slot, its insertion indices follow the
// current ones, and its particles go after the current slots (nothing
// already on the device moves).
tmpm_status commit_staged(tmpm_session* s, tmpm_staged_scene& staged, uint32_t* scene_out) {
  SceneState& scene = staged.scene;
  uint32_t slot = 0;
  while (slot < s->scenes.size() && s->scenes[slot].active) ++slot;
  if (slot >= 0xffffu) return fail(TMPM_CAPACITY, "too many scenes");
  if (!s->prepared) {
    std::vector<SceneState> previous = s->scenes;
    if (slot == s->scenes.size()) s->scenes.push_back(scene);
    else s->scenes[slot] = scene;
    tmpm_status st = reprepare(s, previous);
    if (st != TMPM_OK) return st;
    if (scene_out) *scene_out = slot;
    return TMPM_OK;
  }
  if ((uint64_t)s->id_space + scene.world.size() >= (1ull << 31)) return fail(TMPM_CAPACITY, "too many particles");
  // rebase the staged packing: insertion indices and the scene index of every class
  const uint32_t first = s->id_space;
  scene.first = first;
  PackedParticles& pk = *staged.pk;
  for (size_t d = 0; d < pk.disabled.size(); d += 2)
    pk.disabled[d].w = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, pk.disabled[d].w) + first);
  // the staged classes take the slot's scene index and are interned in the
  // order the particles meet them; the particles then map to the ids
  std::vector<uint32_t> ids(staged.class_table.size());
  for (size_t c = 0; c < staged.class_table.size(); ++c) {
    DevClass dc = staged.class_table[c];
    dc.region |= slot << 16;
    ids[c] = intern_class(s, dc);
  }
  const std::vector<uint32_t>& local = staged.class_local;
  parallel_for(pk.size(), [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      pk.cls[i] = ids[local[i]];
      pk.ins[i] += first;
    }
  });
  if (slot == s->scenes.size()) s->scenes.push_back(std::move(scene));
  else s->scenes[slot] = std::move(scene);
  s->plastic = classes_plastic(s);
  // from here on device state changes; a failure invalidates the batch
  auto broken = [&](tmpm_status code) {
    s->valid = false;
    return code;
  };
  tmpm_status st = TMPM_OK;
  const uint64_t need = (uint64_t)s->n + pk.size();
  if (need > s->cap) {
    const uint32_t cap = (uint32_t)std::max<uint64_t>(need + need / 4, (uint64_t)s->cap + s->cap / 2);
    ParticleBuffers grown, spare, synced;
    if (alloc_particles(grown, cap) != TMPM_OK || alloc_particles(spare, cap) != TMPM_OK ||
        alloc_particles(synced, cap) != TMPM_OK)
      return broken(TMPM_CAPACITY);
    if (copy_particles(grown, cap, s->buf[s->cur], s->cap, s->n, s->stream) != TMPM_OK ||
        copy_particles(synced, cap, s->synced, s->cap, s->synced_n, s->stream) != TMPM_OK)
      return broken(TMPM_READBACK);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return broken(fail(TMPM_READBACK, "device copy failed"));
    auto swap_bufs = [](ParticleBuffers& a, ParticleBuffers& b) {
      std::swap(a.pos_mass.ptr, b.pos_mass.ptr); std::swap(a.pos_mass.bytes, b.pos_mass.bytes);
      std::swap(a.vel_vol.ptr, b.vel_vol.ptr); std::swap(a.vel_vol.bytes, b.vel_vol.bytes);
      std::swap(a.C.ptr, b.C.ptr); std::swap(a.C.bytes, b.C.bytes);
      std::swap(a.F.ptr, b.F.ptr); std::swap(a.F.bytes, b.F.bytes);
      std::swap(a.P.ptr, b.P.ptr); std::swap(a.P.bytes, b.P.bytes);
      std::swap(a.cls.ptr, b.cls.ptr); std::swap(a.cls.bytes, b.cls.bytes);
      std::swap(a.ins_id.ptr, b.ins_id.ptr); std::swap(a.ins_id.bytes, b.ins_id.bytes);
    };
    swap_bufs(s->buf[s->cur], grown);
    swap_bufs(s->buf[1 - s->cur], spare);
    swap_bufs(s->synced, synced);
    s->cap = cap;
  }
  if ((st = upload_particles(s, pk, s->n)) != TMPM_OK) return broken(st);
  s->n += (uint32_t)pk.size();
  s->n_enabled = s->n;
  s->id_space += s->scenes[slot].count;
  if ((st = ensure_force_capacity(s, s->id_space + s->id_space / 4)) != TMPM_OK) return broken(st);
  if ((st = upload_scene_forces(s, slot)) != TMPM_OK) return broken(st);
  s->disabled_host.insert(s->disabled_host.end(), pk.disabled.begin(), pk.disabled.end());
  if ((st = upload_disabled(s)) != TMPM_OK) return broken(st);
  if ((st = upload_classes(s)) != TMPM_OK) return broken(st);
  if ((st = upload_rigid(s)) != TMPM_OK) return broken(st);
  if ((st = upload_obstacles(s)) != TMPM_OK) return broken(st);
  if ((st = upload_scene(s)) != TMPM_OK) return broken(st);
  // topology: room for the blocks this scene adds (and the per-particle arrays)
  s->last_pblocks += (uint32_t)staged.pblocks;
  s->last_gblocks += (uint32_t)staged.gblocks;
  const Topo& t = s->topo;
  if (2ull * s->last_pblocks > t.pblock_cap || 2ull * s->last_gblocks > t.grid_cap || s->topo_cap < s->cap) {
    const uint64_t gcap = std::min<uint64_t>(std::max<uint64_t>(t.grid_cap, 4ull * s->last_gblocks), 1ull << 26);
    const uint64_t pcap = std::min<uint64_t>(std::max<uint64_t>(t.pblock_cap, 4ull * s->last_pblocks), s->cap);
    if ((st = alloc_topology(s, (uint32_t)pcap, (uint32_t)gcap)) != TMPM_OK) return broken(st);
  }
  if (scene_out) *scene_out = slot;
  return TMPM_OK;
}

tmpm_status set_scene_force(tmpm_session* s, uint32_t k, uint64_t index, const double force[3]) {
  SceneState& scene = s->scenes[k];
  if (!scene.stepping()) return fail(scene.status == TMPM_OK ? TMPM_INVALID_INPUT : scene.status, "scene is stopped");
  if (!force) return fail(TMPM_INVALID_INPUT, "null force");
  if (index >= scene.world.size()) return fail(TMPM_INVALID_INPUT, "particle index out of range");
  float f[4];
  if (!finite_f32(force[0], f[0]) || !finite_f32(force[1], f[1]) || !finite_f32(force[2], f[2]) ||
      !finite_f32(scene.world.damping[index], f[3]))
    return fail(TMPM_INVALID_INPUT, "force not representable in f32");
  if (s->prepared) {
    float4 v = make_float4(f[0], f[1], f[2], 0.f);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->force.as<float4>() + scene.first + index, &v, sizeof(v), hipMemcpyHostToDevice));
    const bool was = scene.force_pending;
    scene.force_pending = true;
    if (!was) {
      tmpm_status st = upload_scene(s);
      if (st != TMPM_OK) return st;
    }
  }
  for (int q = 0; q < 3; ++q) scene.world.force[3 * index + q] = force[q];
  return TMPM_OK;
}

tmpm_status set_scene_obstacles(tmpm_session* s, uint32_t k, const tmpm_obstacle* obs, uint32_t count) {
  SceneState& scene = s->scenes[k];
  if (!scene.stepping()) return fail(scene.status == TMPM_OK ? TMPM_INVALID_INPUT : scene.status, "scene is stopped");
  if (count > 0 && !obs) return fail(TMPM_INVALID_INPUT, "null obstacles");
  for (uint32_t i = 0; i < count; ++i)
    if (!obstacle_valid(obs[i])) return fail(TMPM_INVALID_INPUT, "invalid obstacle");
  std::vector<tmpm_obstacle> backup;
  std::vector<std::vector<double>> backup_planes;
  for (auto& o : scene.world.obstacles) {
    backup.push_back(o.o);
    backup_planes.push_back(o.planes);
  }
  for (size_t i = 0; i < backup.size(); ++i) backup[i].planes = backup_planes[i].empty() ? nullptr : backup_planes[i].data();
  scene.world.set_obstacles(obs, count);
  if (s->prepared) {
    tmpm_status st = upload_obstacles(s);
    if (st == TMPM_OK) st = upload_scene(s);
    if (st != TMPM_OK) {
      std::string msg = g_last_error;
      scene.world.set_obstacles(backup.data(), (uint32_t)backup.size());
      if (upload_obstacles(s) != TMPM_OK || upload_scene(s) != TMPM_OK) s->valid = false;
      g_last_error = msg;
      return st;
    }
  }
  return TMPM_OK;
}

}  // namespace

extern "C" {

tmpm_status tmpm_session_create(tmpm_context* ctx, const tmpm_world_desc* wd, double substep_dt, tmpm_p2g_mode mode,
                                tmpm_session** out) {
  try {
    if (!wd) return fail(TMPM_INVALID_INPUT, "null argument");
    return create_session(ctx, wd, &substep_dt, 1, mode, false, out);
  } catch (...) {
    return exception_status();
  }
}

void tmpm_session_destroy(tmpm_session* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  collect_timers(s);
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  destroy_graphs(s);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

tmpm_status tmpm_session_submit_steps(tmpm_session* s, uint32_t steps) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
    if (steps == 0 || steps > 64) return fail(TMPM_INVALID_INPUT, "steps must be within 1..=64");
    if (s->pending + steps < s->pending) return fail(TMPM_CAPACITY, "pending substep count overflow");
    for (auto& scene : s->scenes) {
      if (!scene.stepping()) continue;
      tmpm_status st = validate_resident_world(scene);
      if (st != TMPM_OK) return st;
    }
    if (s->prepared) {
      if (hipSetDevice(s->ctx->device) != hipSuccess) return fail(TMPM_READBACK, "cannot select device");
      if (s->audit_on) {
        tmpm_status a = reserve_audit(s, steps);
        if (a != TMPM_OK) return a;
      }
      for (uint32_t k = 0; k < steps; ++k) {
        bool force = false;
        for (auto& scene : s->scenes) {
          force = force || scene.force_pending;
          scene.force_pending = false;
        }
        if (s->audit_on)
          s->audit_next = s->audit_dev.as<double>() + (s->audit_used++) * (size_t)s->audit_stride * kAuditFields;
        hipError_t e = enqueue_substep(s, force);
        if (e != hipSuccess) {
          s->valid = false;
          return fail(TMPM_READBACK, std::string("kernel launch failed: ") + hipGetErrorString(e));
        }
      }
    }
    s->pending += steps;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_wait(tmpm_session* s) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_synchronize(tmpm_session* s) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
    if (s->pending == 0) return TMPM_OK;
    const uint64_t steps = s->pending;
    s->pending = 0;
    const size_t ns = s->scenes.size();
    std::vector<uint32_t> scene_bits(ns, 0);
    uint32_t counters[CNT_COUNT] = {};
    if (s->prepared) {
      hipError_t e = hipStreamSynchronize(s->stream);
      if (e != hipSuccess) {
        s->valid = false;
        return fail(TMPM_READBACK, std::string("device failure: ") + hipGetErrorString(e));
      }
      collect_timers(s);
      if (s->audit_used && collect_audit(s) != TMPM_OK) {
        s->valid = false;
        return TMPM_READBACK;
      }
      if (hipMemcpy(counters, s->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(scene_bits.data(), s->scene_err.ptr, 4 * ns, hipMemcpyDeviceToHost) != hipSuccess) {
        s->valid = false;
        return fail(TMPM_READBACK, "counter readback failed");
      }
      // Capacity and internal errors concern every scene: the session fails.
      // A single-scene session also fails on its scene's errors; a batch stops
      // only the scenes concerned.
      const uint32_t global = s->batch ? (counters[CNT_ERROR] & (ERR_TOPOLOGY | ERR_INTERNAL)) : counters[CNT_ERROR];
      tmpm_status st = status_from_error_bits(global);
      if (st != TMPM_OK) {
        s->valid = false;
        return st;
      }
      s->last_pblocks = counters[CNT_PARTICLE_BLOCKS];
      s->last_gblocks = counters[CNT_GRID_BLOCKS];
      if (counters[CNT_REACT_SATURATED]) {
        s->react_saturated += counters[CNT_REACT_SATURATED];
        const uint32_t zero = 0;
        if (hipMemcpy(s->counters.as<uint32_t>() + CNT_REACT_SATURATED, &zero, 4, hipMemcpyHostToDevice) !=
            hipSuccess) {
          s->valid = false;
          return fail(TMPM_READBACK, "counter reset failed");
        }
      }
    }
    std::vector<char> commit(ns, 0);
    tmpm_status st = TMPM_OK;
    bool coupling_read = false;
    if (!s->batch) {
      // eager: the world is read back at every synchronize
      std::vector<HostWorld> updated;
      std::vector<tmpm_status> scene_st;
      st = read_back(s, updated, scene_st);
      if (st != TMPM_OK) {
        s->valid = false;
        return st;
      }
      if (scene_st[0] != TMPM_OK) {
        s->valid = false;
        return fail(scene_st[0], "MPM integration produced a non-finite state");
      }
      SceneState& scene = s->scenes[0];
      const uint64_t before = scene.world.substeps;
      updated[0].substeps = before + steps < before ? UINT64_MAX : before + steps;
      scene.set_world(updated[0]);
      commit[0] = 1;
    } else {
      // lazy: stopped scenes take their state of the previous synchronize from
      // the device copy; the others only advance their substep counts and are
      // read on demand from the new copy
      bool stopped = false;
      for (size_t k = 0; k < ns; ++k) {
        SceneState& scene = s->scenes[k];
        if (!scene.stepping()) continue;
        if (scene_bits[k] == 0) {
          const uint64_t before = scene.world.substeps;
          scene.world.substeps = before + steps < before ? UINT64_MAX : before + steps;
          if (s->prepared && scene.count) scene.device_newer = true;
          commit[k] = 1;
          continue;
        }
        st = materialize(s, (uint32_t)k);
        if (st != TMPM_OK) {
          s->valid = false;
          return st;
        }
        scene.status = status_from_error_bits(scene_bits[k]);  // the last synchronized world stays readable
        stopped = true;
        if (s->prepared && scene.count) {
          launch_retire(s->parts(s->cur), scene.first, scene.count, s->stream);
          HIP_TRY(hipGetLastError());
        }
      }
      // The committed scenes' rigid bodies, obstacle poses and reactions are
      // read first: a stopped scene's uploads below rewrite every scene's
      // bodies and obstacles from the host copies.
      st = read_coupling(s, commit);
      if (st != TMPM_OK) {
        s->valid = false;
        return st;
      }
      coupling_read = true;
      if (stopped) {
        s->substeps_since_reorder = s->reorder_interval;  // the next substep drops the retired slots
        st = upload_rigid(s);
        if (st == TMPM_OK) st = upload_obstacles(s);
        if (st == TMPM_OK) st = upload_scene(s);
        if (st != TMPM_OK) {
          s->valid = false;
          return st;
        }
        HIP_TRY(hipMemset(s->scene_err.ptr, 0, 4ull * s->scene_cap));
        uint32_t err = counters[CNT_ERROR] & ~(ERR_RESIDENT | ERR_FIXED_RANGE);
        HIP_TRY(hipMemcpy(s->counters.as<uint32_t>() + CNT_ERROR, &err, 4, hipMemcpyHostToDevice));
      }
      if (s->prepared) {
        if (s->compacted_last && !stopped) s->n = std::min<uint32_t>(s->n, counters[CNT_BINNED]);
        st = copy_particles(s->synced, s->cap, s->buf[s->cur], s->cap, s->n, s->stream);
        if (st != TMPM_OK) {
          s->valid = false;
          return st;
        }
        s->synced_n = s->n;
      }
    }
    if (!coupling_read) st = read_coupling(s, commit);
    if (st != TMPM_OK) {
      s->valid = false;
      return st;
    }
    if (s->prepared) {
      st = maybe_grow_topology(s, counters);
      if (st != TMPM_OK) {
        s->valid = false;
        return st;
      }
    }
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

// gpu.rs:413-450 (synchronize_async): a poll instead of an await.  While the
// device still runs the submitted substeps it returns TMPM_OK with *done_out
// = 0 and changes nothing (the pending substeps stay pending, so the caller
// may stop polling and synchronize later: the reference's cancellation keeps
// them too).  Once they have finished it validates and commits exactly as
// tmpm_session_synchronize does and sets *done_out = 1, also when that
// fails (the failure invalidates the session as there).
tmpm_status tmpm_session_synchronize_async(tmpm_session* s, int32_t* done_out) {
  try {
    if (!s || !done_out) return fail(TMPM_INVALID_INPUT, "null argument");
    *done_out = 0;
    if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
    if (s->pending == 0) {
      *done_out = 1;
      return TMPM_OK;
    }
    if (s->prepared) {
      const hipError_t q = hipStreamQuery(s->stream);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // not a failure: keep it out of the next launch check
        return TMPM_OK;
      }
      if (q != hipSuccess) {
        s->valid = false;
        *done_out = 1;
        return fail(TMPM_READBACK, std::string("device failure: ") + hipGetErrorString(q));
      }
    }
    *done_out = 1;
    return tmpm_session_synchronize(s);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_step(tmpm_session* s, uint32_t steps) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
    tmpm_status st = tmpm_session_submit_steps(s, steps);
    if (st != TMPM_OK) return st;
    return tmpm_session_synchronize(s);
  } catch (...) {
    return api_exception(s);
  }
}

// gpu.rs:2208-2240: CFL-limited substeps, each validated on the host.
tmpm_status tmpm_session_step_adaptive(tmpm_session* s, double dt) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    if (s->batch) return fail(TMPM_INVALID_INPUT, "adaptive stepping takes a single-scene session");
    if (!std::isfinite(dt) || dt <= 0.0) return fail(TMPM_INVALID_INPUT, "dt must be positive");
    SceneState& scene = s->scenes[0];
    for (auto& o : scene.world.obstacles)
      if (!obstacle_valid(o.o)) return fail(TMPM_INVALID_INPUT, "invalid obstacle");
    double remaining = dt;
    size_t count = 0;
    const double saved_dt = scene.substep_dt;
    while (remaining > 0.0) {
      double stable = scene.stable_timestep();
      if (!std::isfinite(stable) || stable <= 1e-12 || count >= 100000)
        return fail(TMPM_EXCESSIVE_SUBSTEPS, "MPM timestep is too small for the requested duration");
      double sub = std::min(remaining, stable);
      if (sub != scene.substep_dt) {
        scene.substep_dt = sub;
        st = build_scene(scene);
        if (st == TMPM_OK && s->prepared) st = upload_obstacles(s);
        if (st == TMPM_OK && s->prepared) st = upload_scene(s);
        if (st != TMPM_OK) {
          scene.substep_dt = saved_dt;
          (void)build_scene(scene);
          return st;
        }
      }
      st = tmpm_session_step(s, 1);
      if (st != TMPM_OK) return st;
      remaining = std::max(remaining - sub, 0.0);
      ++count;
    }
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

uint64_t tmpm_session_pending_substeps(const tmpm_session* s) { return s ? s->pending : 0; }
uint64_t tmpm_session_reaction_saturations(const tmpm_session* s) { return s ? s->react_saturated : 0; }
tmpm_status tmpm_session_memory(const tmpm_session* s, uint64_t* device_bytes, uint64_t* particle_capacity) {
  if (!s || !device_bytes || !particle_capacity) return fail(TMPM_INVALID_INPUT, "null argument");
  *device_bytes = s->device_bytes();
  *particle_capacity = s->cap;
  return TMPM_OK;
}
int32_t tmpm_session_valid(const tmpm_session* s) { return s && s->valid ? 1 : 0; }
uint64_t tmpm_session_substeps(const tmpm_session* s) { return s && !s->scenes.empty() ? s->world().substeps : 0; }
uint64_t tmpm_session_particle_count(const tmpm_session* s) { return s && !s->scenes.empty() ? s->world().size() : 0; }

tmpm_status tmpm_session_read_particles(const tmpm_session* s, tmpm_particles* out) {
  try {
    if (!s || !out || s->scenes.empty()) return fail(TMPM_INVALID_INPUT, "null argument");
    if (out->count < s->world().size()) return fail(TMPM_CAPACITY, "output arrays too small");
    return copy_particles_out(s->world(), out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_particle_force(tmpm_session* s, uint64_t index, const double force[3]) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    return set_scene_force(s, 0, index, force);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_obstacles(tmpm_session* s, const tmpm_obstacle* obs, uint32_t count) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    return set_scene_obstacles(s, 0, obs, count);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_add_particles(tmpm_session* s, const tmpm_particles* particles, const tmpm_material* materials,
                                       uint32_t material_count, uint64_t* chunk_out) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    if (!particles) return fail(TMPM_INVALID_INPUT, "null argument");
    if (particles->count == 0 || !validate_particles(*particles, materials, material_count))
      return fail(TMPM_INVALID_INPUT, "invalid MPM particle batch");
    SceneState& scene = s->scenes[0];
    if (scene.world.next_chunk_id == UINT64_MAX) return fail(TMPM_PARTICLE_CAPACITY, "chunk ids exhausted");
    std::vector<SceneState> previous = s->scenes;
    HostWorld updated = scene.world;
    const uint64_t chunk = updated.next_chunk_id++;
    updated.append(*particles, materials, chunk);
    SceneState trial = scene;
    trial.set_world(updated);
    st = validate_resident_world(trial);
    if (st != TMPM_OK) return st;
    scene = trial;
    st = reprepare(s, previous);
    if (st != TMPM_OK) return st;
    if (chunk_out) *chunk_out = chunk;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_remove_chunk(tmpm_session* s, uint64_t chunk, uint64_t* removed_out) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    SceneState& scene = s->scenes[0];
    std::vector<SceneState> previous = s->scenes;
    HostWorld updated = scene.world;
    const size_t removed = updated.remove_chunk(chunk);
    if (removed == 0) return fail(TMPM_UNKNOWN_CHUNK, "MPM particle chunk was not found");
    SceneState trial = scene;
    trial.set_world(updated);
    st = validate_resident_world(trial);
    if (st != TMPM_OK) return st;
    scene = trial;
    st = reprepare(s, previous);
    if (st != TMPM_OK) return st;
    if (removed_out) *removed_out = removed;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

// ---------------------------------------------------------------------------
// Batches: several independent scenes stepped together
// ---------------------------------------------------------------------------
tmpm_status tmpm_batch_create(tmpm_context* ctx, const tmpm_world_desc* worlds, const double* substep_dt,
                              uint32_t scene_count, tmpm_p2g_mode mode, tmpm_session** out) {
  try {
    return create_session(ctx, worlds, substep_dt, scene_count, mode, true, out);
  } catch (...) {
    return exception_status();
  }
}

uint32_t tmpm_batch_scene_count(const tmpm_session* s) { return s ? (uint32_t)s->scenes.size() : 0; }

tmpm_status tmpm_batch_scene_status(const tmpm_session* s, uint32_t scene) {
  try {
    tmpm_status st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return s->scenes[scene].status;
  } catch (...) {
    return api_exception(s);
  }
}

uint64_t tmpm_batch_scene_particle_count(const tmpm_session* s, uint32_t scene) {
  return check_scene(s, scene) == TMPM_OK ? s->scenes[scene].world.size() : 0;
}

uint64_t tmpm_batch_scene_substeps(const tmpm_session* s, uint32_t scene) {
  return check_scene(s, scene) == TMPM_OK ? s->scenes[scene].world.substeps : 0;
}

tmpm_status tmpm_batch_read_particles(const tmpm_session* s, uint32_t scene, tmpm_particles* out) {
  try {
    tmpm_status st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    if (!out) return fail(TMPM_INVALID_INPUT, "null argument");
    if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
    // the world of the last synchronize, read from the device on first use
    st = materialize(const_cast<tmpm_session*>(s), scene);
    if (st != TMPM_OK) return st;
    const HostWorld& w = s->scenes[scene].world;
    if (out->count < w.size()) return fail(TMPM_CAPACITY, "output arrays too small");
    return copy_particles_out(w, out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_set_particle_force(tmpm_session* s, uint32_t scene, uint64_t index, const double force[3]) {
  try {
    tmpm_status st = check_editable(s);
    if (st == TMPM_OK) st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return set_scene_force(s, scene, index, force);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_set_obstacles(tmpm_session* s, uint32_t scene, const tmpm_obstacle* obs, uint32_t count) {
  try {
    tmpm_status st = check_editable(s);
    if (st == TMPM_OK) st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return set_scene_obstacles(s, scene, obs, count);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_stage_scene(const tmpm_world_desc* world, double substep_dt, tmpm_staged_scene** out) {
  try {
    if (!out) return fail(TMPM_INVALID_INPUT, "null argument");
    *out = nullptr;
    auto staged = std::make_unique<tmpm_staged_scene>();
    tmpm_status st = stage_scene(world, substep_dt, nullptr, *staged);
    if (st != TMPM_OK) return st;
    *out = staged.release();
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

void tmpm_staged_scene_destroy(tmpm_staged_scene* staged) {
  if (!staged) return;
  pack_pool().release(std::move(staged->pk));
  delete staged;
}

tmpm_status tmpm_batch_add_staged(tmpm_session* s, tmpm_staged_scene* staged, uint32_t* scene_out) {
  try {
    if (!staged) return fail(TMPM_INVALID_INPUT, "null staged scene");
    std::unique_ptr<tmpm_staged_scene, void (*)(tmpm_staged_scene*)> own(staged, tmpm_staged_scene_destroy);
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    if (!s->batch) return fail(TMPM_INVALID_INPUT, "not a batch");
    return commit_staged(s, *staged, scene_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_add_scene(tmpm_session* s, const tmpm_world_desc* world, double substep_dt,
                                 uint32_t* scene_out) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    if (!s->batch) return fail(TMPM_INVALID_INPUT, "not a batch");
    tmpm_staged_scene staged;
    // reuse the best-fitting emptied world of a retired scene, if one is large enough
    const size_t need = world ? 3 * (size_t)world->particles.count : 0;
    size_t best = s->spare_worlds.size();
    for (size_t i = 0; i < s->spare_worlds.size(); ++i)
      if (s->spare_worlds[i].pos.capacity() >= need &&
          (best == s->spare_worlds.size() || s->spare_worlds[i].pos.capacity() < s->spare_worlds[best].pos.capacity()))
        best = i;
    HostWorld spare;
    if (best < s->spare_worlds.size()) {
      spare = std::move(s->spare_worlds[best]);
      s->spare_worlds.erase(s->spare_worlds.begin() + best);
    }
    st = stage_scene(world, substep_dt, &spare, staged);
    if (st == TMPM_OK) st = commit_staged(s, staged, scene_out);
    pack_pool().release(std::move(staged.pk));
    return st;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_remove_scene(tmpm_session* s, uint32_t scene) {
  try {
    tmpm_status st = check_editable(s);
    if (st == TMPM_OK) st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    if (!s->batch) return fail(TMPM_INVALID_INPUT, "not a batch");
    SceneState& gone = s->scenes[scene];
    const uint32_t first = gone.first, count = gone.count;
    if (s->prepared && count > 0) {
      // Retirement on the device: the scene's particles lose their mass (the
      // next reorder drops them) and the insertion indices above its range
      // close the gap, here and in the synchronized copy.
      Parts p = s->parts(s->cur);
      launch_retire(p, first, count, s->stream);
      launch_renumber(p.ins_id, s->n, first, count, s->stream);
      launch_renumber(s->synced.ins_id.as<uint32_t>(), s->synced_n, first, count, s->stream);
      HIP_TRY(hipGetLastError());
      const uint32_t above = s->id_space - (first + count);
      if (above > 0) {
        if (s->staging.bytes < 16ull * above) HIP_TRY(s->staging.alloc(16ull * above));
        HIP_TRY(hipMemcpyAsync(s->staging.ptr, s->force.as<float4>() + first + count, 16ull * above,
                               hipMemcpyDeviceToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->force.as<float4>() + first, s->staging.ptr, 16ull * above, hipMemcpyDeviceToDevice,
                               s->stream));
      }
      HIP_TRY(hipStreamSynchronize(s->stream));
      s->substeps_since_reorder = s->reorder_interval;
    }
    for (auto& other : s->scenes)
      if (other.active && &other != &gone && other.first > first) other.first -= count;
    s->id_space -= count;
    // snapshot records of disabled particles: drop the scene's, shift the others
    std::vector<float4> kept;
    for (size_t i = 0; i + 1 < s->disabled_host.size(); i += 2) {
      float4 a = s->disabled_host[i];
      uint32_t g = __builtin_bit_cast(uint32_t, a.w);
      if (g - first < count) continue;
      if (g >= first + count) a.w = __builtin_bit_cast(float, g - count);
      kept.push_back(a);
      kept.push_back(s->disabled_host[i + 1]);
    }
    s->disabled_host.swap(kept);
    {
      // keep the emptied world's memory for a later admission (at most 8
      // worlds and 2 GiB)
      HostWorld w = std::move(gone.world);
      w.clear_keep();
      size_t kept_bytes = w.capacity_bytes();
      for (const HostWorld& o : s->spare_worlds) kept_bytes += o.capacity_bytes();
      if (s->spare_worlds.size() < 8 && kept_bytes <= (2ull << 30)) s->spare_worlds.push_back(std::move(w));
    }
    s->scenes[scene] = SceneState();
    s->scenes[scene].active = false;
    if (!s->prepared) return TMPM_OK;
    if ((st = upload_disabled(s)) != TMPM_OK || (st = upload_rigid(s)) != TMPM_OK ||
        (st = upload_obstacles(s)) != TMPM_OK || (st = upload_scene(s)) != TMPM_OK) {
      s->valid = false;
      return st;
    }
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_reactions(tmpm_session* s, int32_t enabled) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    const bool was = s->reactions_on;
    s->reactions_on = enabled != 0;
    if (!s->prepared || was == s->reactions_on) return TMPM_OK;
    st = upload_obstacles(s);
    if (st == TMPM_OK) st = upload_scene(s);
    if (st != TMPM_OK) {
      std::string msg = g_last_error;
      s->reactions_on = was;
      if (upload_obstacles(s) != TMPM_OK || upload_scene(s) != TMPM_OK) s->valid = false;
      g_last_error = msg;
    }
    return st;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_read_reactions(const tmpm_session* s, double* out, uint32_t cap, uint32_t* count_out) {
  try {
    if (!s || s->scenes.empty()) return fail(TMPM_INVALID_INPUT, "null session");
    return read_scene_reactions(s->scenes[0], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_rigid_bodies(tmpm_session* s, const tmpm_rigid_body* bodies, uint32_t body_count,
                                          const tmpm_rigid_mapping* mappings, uint32_t mapping_count,
                                          const double gravity[3], double max_linear_speed, int32_t integrate) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    return set_scene_rigid(s, 0, bodies, body_count, mappings, mapping_count, gravity, max_linear_speed, integrate);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_read_rigid_bodies(const tmpm_session* s, tmpm_rigid_body* out, uint32_t cap,
                                           uint32_t* count_out) {
  try {
    if (!s || s->scenes.empty()) return fail(TMPM_INVALID_INPUT, "null session");
    return read_scene_bodies(s->scenes[0], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_read_obstacles(const tmpm_session* s, tmpm_obstacle* out, uint32_t cap, uint32_t* count_out) {
  try {
    if (!s || s->scenes.empty()) return fail(TMPM_INVALID_INPUT, "null session");
    return read_scene_obstacles(s->scenes[0], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_read_reactions(const tmpm_session* s, uint32_t scene, double* out, uint32_t cap,
                                      uint32_t* count_out) {
  try {
    tmpm_status st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return read_scene_reactions(s->scenes[scene], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_set_rigid_bodies(tmpm_session* s, uint32_t scene, const tmpm_rigid_body* bodies,
                                        uint32_t body_count, const tmpm_rigid_mapping* mappings,
                                        uint32_t mapping_count, const double gravity[3], double max_linear_speed,
                                        int32_t integrate) {
  try {
    tmpm_status st = check_editable(s);
    if (st == TMPM_OK) st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return set_scene_rigid(s, scene, bodies, body_count, mappings, mapping_count, gravity, max_linear_speed, integrate);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_read_rigid_bodies(const tmpm_session* s, uint32_t scene, tmpm_rigid_body* out, uint32_t cap,
                                         uint32_t* count_out) {
  try {
    tmpm_status st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return read_scene_bodies(s->scenes[scene], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_batch_read_obstacles(const tmpm_session* s, uint32_t scene, tmpm_obstacle* out, uint32_t cap,
                                      uint32_t* count_out) {
  try {
    tmpm_status st = check_scene(s, scene);
    if (st != TMPM_OK) return st;
    return read_scene_obstacles(s->scenes[scene], out, cap, count_out);
  } catch (...) {
    return api_exception(s);
  }
}


tmpm_status tmpm_session_encode_snapshot(tmpm_session* s, void* device_output, uint64_t output_bytes, void* stream,
                                         uint32_t* count_out) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
    if (count_out) *count_out = 0;
    if (!s->prepared) return TMPM_OK;
    uint64_t bytes = 32ull * s->id_space;
    if (!device_output || output_bytes < bytes) return fail(TMPM_CAPACITY, "snapshot buffer too small");
    hipStream_t target = stream ? static_cast<hipStream_t>(stream) : s->stream;
    if (target != s->stream) {
      hipEvent_t ev;
      HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      HIP_TRY(hipEventRecord(ev, s->stream));
      HIP_TRY(hipStreamWaitEvent(target, ev, 0));
      HIP_TRY(hipEventDestroy(ev));
    }
    launch_snapshot(s->parts(s->cur), static_cast<float4*>(device_output), target);
    launch_snapshot_disabled(s->disabled_rec.as<float4>(), (uint32_t)(s->disabled_host.size() / 2),
                             static_cast<float4*>(device_output), target);
    HIP_TRY(hipGetLastError());
    if (target != s->stream) {
      hipEvent_t ev;
      HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      HIP_TRY(hipEventRecord(ev, target));
      HIP_TRY(hipStreamWaitEvent(s->stream, ev, 0));
      HIP_TRY(hipEventDestroy(ev));
    }
    if (count_out) *count_out = s->id_space;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

double tmpm_session_stable_timestep(const tmpm_session* s) {
  return s && !s->scenes.empty() ? s->world().stable_timestep() : 0.0;
}

tmpm_status tmpm_session_topology_stats(tmpm_session* s, tmpm_topology_stats* out) {
  try {
    if (!s || !out) return fail(TMPM_INVALID_INPUT, "null argument");
    if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
    std::memset(out, 0, sizeof(*out));
    std::vector<uint32_t> c;
    tmpm_status st = run_diagnostic_binning(s, c);
    if (st != TMPM_OK) return st;
    out->live_nodes = c[CNT_LIVE_NODES];
    out->grid_blocks = c[CNT_GRID_BLOCKS];
    out->particle_blocks = c[CNT_PARTICLE_BLOCKS];
    out->binned_particles = c[CNT_BINNED];
    return status_from_error_bits(c[CNT_ERROR] & ~ERR_RESIDENT);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_export_nodes(tmpm_session* s, int32_t* keys, uint64_t cap, uint64_t* count_out) {
  try {
    if (!s || !count_out) return fail(TMPM_INVALID_INPUT, "null argument");
    if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
    std::vector<uint32_t> c;
    tmpm_status st = run_diagnostic_binning(s, c);
    if (st != TMPM_OK) return st;
    *count_out = 0;
    if (!s->prepared) return TMPM_OK;
    uint32_t ng = c[CNT_GRID_BLOCKS];
    std::vector<int4> gk(ng);
    std::vector<uint32_t> masks(2ull * ng);
    HIP_TRY(hipMemcpy(gk.data(), s->grid_key.ptr, 16ull * ng, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(masks.data(), s->node_mask.ptr, 8ull * ng, hipMemcpyDeviceToHost));
    std::vector<std::array<int32_t, 4>> nodes;
    for (uint32_t g = 0; g < ng; ++g)
      for (int l = 0; l < 64; ++l)
        if (masks[2 * g + (l >> 5)] & (1u << (l & 31)))
          nodes.push_back({gk[g].x * 4 + (l >> 4), gk[g].y * 4 + ((l >> 2) & 3), gk[g].z * 4 + (l & 3),
                           (int32_t)((uint32_t)gk[g].w & 0xffffu)});
    std::sort(nodes.begin(), nodes.end());
    for (uint64_t i = 0; i < nodes.size() && i < cap; ++i)
      for (int k = 0; k < 4; ++k) keys[4 * i + k] = nodes[i][k];
    *count_out = nodes.size();
    return status_from_error_bits(c[CNT_ERROR] & ~ERR_RESIDENT);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_export_blocks(tmpm_session* s, int32_t* block_keys, uint64_t cap, int64_t* particle_block,
                                       uint64_t* count_out) {
  try {
    if (!s || !count_out) return fail(TMPM_INVALID_INPUT, "null argument");
    if (s->pending) return fail(TMPM_PENDING_STEPS, "steps are pending; synchronize first");
    std::vector<uint32_t> c;
    tmpm_status st = run_diagnostic_binning(s, c);
    if (st != TMPM_OK) return st;
    *count_out = 0;
    size_t total = s->scenes.empty() ? 0 : s->world().size();
    if (particle_block)
      for (size_t i = 0; i < total; ++i) particle_block[i] = -1;
    if (!s->prepared) return TMPM_OK;
    uint32_t ng = c[CNT_GRID_BLOCKS];
    uint32_t nb = s->topo.bucket_mask + 1;
    std::vector<int4> gk(ng);
    std::vector<uint32_t> pslot(s->n), bslot(nb), bpb(nb), ins(s->n);
    HIP_TRY(hipMemcpy(gk.data(), s->grid_key.ptr, 16ull * ng, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pslot.data(), s->pslot.ptr, 4ull * s->n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bslot.data(), s->bucket_slot.ptr, 4ull * nb, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bpb.data(), s->bucket_pblock.ptr, 4ull * nb, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ins.data(), s->buf[s->cur].ins_id.ptr, 4ull * s->n, hipMemcpyDeviceToHost));
    std::vector<std::array<int32_t, 4>> keys;
    std::vector<std::array<int32_t, 4>> pkey(s->n);
    std::vector<char> has(s->n, 0);
    for (uint32_t i = 0; i < s->n; ++i) {
      uint32_t b = pslot[i];
      if (b == kInvalid || bpb[b] == kInvalid || bslot[b] == kInvalid || bslot[b] >= ng) continue;
      int4 k = gk[bslot[b]];
      pkey[i] = {k.x, k.y, k.z, (int32_t)((uint32_t)k.w & 0xffffu)};
      has[i] = 1;
      keys.push_back(pkey[i]);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (uint64_t i = 0; i < keys.size() && i < cap; ++i)
      for (int k = 0; k < 4; ++k) block_keys[4 * i + k] = keys[i][k];
    if (particle_block)
      for (uint32_t i = 0; i < s->n; ++i)
        if (has[i]) particle_block[ins[i]] = std::lower_bound(keys.begin(), keys.end(), pkey[i]) - keys.begin();
    *count_out = keys.size();
    return status_from_error_bits(c[CNT_ERROR] & ~ERR_RESIDENT);
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_timing(tmpm_session* s, int32_t enabled) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timers(s);
    s->kernel_ms.clear();
    s->timing = enabled != 0;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_launch_config(tmpm_session* s, int32_t out[5]) {
  try {
    if (!s || !out) return fail(TMPM_INVALID_INPUT, "null argument");
    uint32_t p2g_threads;
    int p2g_copies;
    effective_p2g_shape(s, p2g_threads, p2g_copies);
    out[0] = (int32_t)p2g_threads;
    out[1] = p2g_copies;
    out[2] = (int32_t)effective_g2p_threads(s);
    out[3] = (int32_t)s->reorder_interval;
    out[4] = s->use_graphs ? 1 : 0;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

// Launch-shape autotuner.  Every candidate shape computes the same results:
// P2G sums integers (order-free), G2P is per particle and sums its reaction
// impulses as rounded integers, and the storage order a reorder produces
// changes no sum.  The trials therefore run real substeps of the simulation, one knob
// at a time (P2G shape, G2P workgroup, reorder interval, graphs), each
// candidate timed over `trial` substeps with events on the session stream.
tmpm_status tmpm_session_autotune(tmpm_session* s, uint32_t trial, uint32_t* substeps_run) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    if (!s->valid) return fail(TMPM_RESIDENT_UNAVAILABLE, "session is invalid after an earlier failure");
    if (trial < 8 || trial > 256) return fail(TMPM_INVALID_INPUT, "trial must be within 8..=256 substeps");
    if (substeps_run) *substeps_run = 0;
    if (!s->prepared || s->n == 0) return TMPM_OK;
    if (s->timing) return fail(TMPM_INVALID_INPUT, "autotuning needs per-kernel timing off");
    // a bucket tuned before (this process, or the TMPM_TUNE_CACHE file): no trials
    const TuneKey key = tune_key(s);
    TuneShape hit;
    if (tune_cache().find(key, hit)) {
      s->p2g_threads = hit.p2g_threads;
      s->p2g_copies = hit.p2g_copies;
      s->threads_pblock = hit.g2p_threads;
      s->reorder_interval = hit.reorder_interval;
      s->use_graphs = hit.graphs != 0;
      s->p2g_shape_fixed = s->g2p_shape_fixed = true;
      if (!s->use_graphs) destroy_graphs(s);
      return TMPM_OK;
    }
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    uint32_t ran = 0;
    tmpm_status st = TMPM_OK;
    // The trials are queued in chunks of at most 64 substeps, the submit
    // granularity.  After each chunk the device counters are read: the
    // topology grows as synchronize would grow it (a spreading scene cannot
    // outgrow its headroom because of tuning), and on any error the tuning
    // stops, so the pending substeps report it at the next synchronize just as
    // untuned stepping would.  Only the chunks' device time is measured.
    bool stopped = false;
    auto run = [&](uint32_t steps, float* ms) -> tmpm_status {
      float total = 0.0f;
      for (uint32_t left = steps; left > 0 && !stopped;) {
        const uint32_t k = std::min<uint32_t>(left, 64);
        if (hipEventRecord(a, s->stream) != hipSuccess) return fail(TMPM_READBACK, "event record failed");
        tmpm_status e = tmpm_session_submit_steps(s, k);
        if (e != TMPM_OK) return e;
        left -= k;
        ran += k;
        if (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess)
          return fail(TMPM_READBACK, "event wait failed");
        float chunk = 0.0f;
        if (hipEventElapsedTime(&chunk, a, b) != hipSuccess) return fail(TMPM_READBACK, "event time failed");
        total += chunk;
        uint32_t counters[CNT_COUNT] = {};
        HIP_TRY(hipMemcpy(counters, s->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost));
        const uint32_t fatal = s->batch ? (counters[CNT_ERROR] & (ERR_TOPOLOGY | ERR_INTERNAL)) : counters[CNT_ERROR];
        if (fatal != 0) {
          stopped = true;
          break;
        }
        e = maybe_grow_topology(s, counters);
        if (e != TMPM_OK) {
          s->valid = false;
          return e;
        }
      }
      if (ms) *ms = total;
      return TMPM_OK;
    };
    // Tries each value of one knob (set by `apply`) and keeps the fastest.
    auto tune = [&](const std::vector<int>& values, const std::function<void(int)>& apply, int current) -> tmpm_status {
      float best_ms = 0.0f;
      int best = current;
      for (size_t i = 0; i < values.size() && !stopped; ++i) {
        apply(values[i]);
        float ms = 0.0f;
        tmpm_status e = run(trial, &ms);
        if (e != TMPM_OK) return e;
        if (stopped) break;
        if (i == 0 || ms < best_ms) {
          best_ms = ms;
          best = values[i];
        }
      }
      apply(best);
      return TMPM_OK;
    };
    st = run(std::min<uint32_t>(trial, 16), nullptr);  // warm-up
    if (st == TMPM_OK && !stopped) {
      effective_p2g_shape(s, s->p2g_threads, s->p2g_copies);  // trials start from the shape in use
      s->p2g_shape_fixed = true;  // the trials' shapes, not the size-based default, run from here on
      const int cur = (int)s->p2g_threads * 10 + s->p2g_copies;
      st = tune({642, 641, 1282, 1281}, [&](int v) { s->p2g_threads = (uint32_t)(v / 10); s->p2g_copies = v % 10; }, cur);
    }
    if (st == TMPM_OK && !stopped) {
      s->threads_pblock = effective_g2p_threads(s);  // trials start from the size in use
      s->g2p_shape_fixed = true;
      st = tune({64, 128, 256}, [&](int v) { s->threads_pblock = (uint32_t)v; }, (int)s->threads_pblock);
    }
    if (st == TMPM_OK && !stopped)
      st = tune({4, 8, 16}, [&](int v) { s->reorder_interval = (uint32_t)v; }, (int)s->reorder_interval);
    if (st == TMPM_OK && !stopped) st = tune({0, 1}, [&](int v) { s->use_graphs = v != 0; }, s->use_graphs ? 1 : 0);
    if (!s->use_graphs) destroy_graphs(s);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (substeps_run) *substeps_run = ran;
    if (st == TMPM_OK && !stopped) {
      // cached under the bucket the session was in when tuning started
      TuneShape v;
      v.p2g_threads = s->p2g_threads;
      v.p2g_copies = s->p2g_copies;
      v.g2p_threads = s->threads_pblock;
      v.reorder_interval = s->reorder_interval;
      v.graphs = s->use_graphs ? 1 : 0;
      tune_cache().put(key, v);
    }
    return st;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_set_audit(tmpm_session* s, int32_t enabled) {
  try {
    tmpm_status st = check_editable(s);
    if (st != TMPM_OK) return st;
    s->audit_on = enabled != 0;
    if (!s->audit_on) {
      s->audit_dev.release();
      s->audit_cap = 0;
    }
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_read_audit(const tmpm_session* s, uint32_t scene, double* out, uint64_t cap_records,
                                    uint64_t* count_out) {
  try {
    if (!s || !count_out || (cap_records && !out)) return fail(TMPM_INVALID_INPUT, "null argument");
    if (scene >= s->scenes.size()) return fail(TMPM_INVALID_INPUT, "no such scene");
    const uint64_t steps = s->audit_host_steps;
    *count_out = steps;
    for (uint64_t k = 0; k < std::min(steps, cap_records); ++k)
      std::memcpy(out + k * kAuditFields,
                  s->audit_host.data() + (k * s->audit_host_stride + scene) * (size_t)kAuditFields,
                  sizeof(double) * kAuditFields);
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

void tmpm_tune_cache_clear(void) { tune_cache().clear(); }
uint64_t tmpm_tune_cache_size(void) { return tune_cache().size(); }

tmpm_status tmpm_session_set_graphs(tmpm_session* s, int32_t enabled) {
  try {
    if (!s) return fail(TMPM_INVALID_INPUT, "null session");
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->use_graphs = enabled != 0;
    if (!s->use_graphs) destroy_graphs(s);
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_session_kernel_times(tmpm_session* s, char* names_csv, uint64_t names_cap, double* avg_ms,
                                      uint64_t* calls, uint32_t cap, uint32_t* count_out) {
  try {
    if (!s || !count_out) return fail(TMPM_INVALID_INPUT, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    collect_timers(s);
    std::string csv;
    uint32_t k = 0;
    for (auto& kv : s->kernel_ms) {
      if (k < cap) {
        if (avg_ms) avg_ms[k] = kv.second.second ? kv.second.first / (double)kv.second.second : 0.0;
        if (calls) calls[k] = kv.second.second;
        if (!csv.empty()) csv += ",";
        csv += kv.first;
      }
      ++k;
    }
    if (names_csv && names_cap) {
      std::strncpy(names_csv, csv.c_str(), names_cap - 1);
      names_csv[names_cap - 1] = '\0';
    }
    *count_out = k;
    return TMPM_OK;
  } catch (...) {
    return api_exception(s);
  }
}

tmpm_status tmpm_sample_closed_mesh_volume(tmpm_context* ctx, const double* vertices, uint64_t vertex_count,
                                           const uint32_t* triangles, uint64_t triangle_count, double spacing,
                                           uint64_t max_candidates, uint64_t max_samples, double* samples_out,
                                           uint64_t* count_out) {
  try {
    if (count_out) *count_out = 0;
    tmpm_status st = mesh_check_triangles(ctx, vertices, vertex_count, triangles, triangle_count, spacing,
                                          max_candidates, max_samples, samples_out, count_out);
    if (st != TMPM_OK) return st;
    // sampling.rs:66-93: every undirected edge used exactly once in each
    // direction (a closed, consistently oriented surface)
    std::map<std::pair<uint32_t, uint32_t>, std::array<uint64_t, 2>> edges;
    for (uint64_t t = 0; t < triangle_count; ++t) {
      const uint32_t* tr = triangles + 3 * t;
      const uint32_t pairs[3][2] = {{tr[0], tr[1]}, {tr[1], tr[2]}, {tr[2], tr[0]}};
      for (const auto& pr : pairs)
        edges[{std::min(pr[0], pr[1]), std::max(pr[0], pr[1])}][pr[0] > pr[1] ? 1 : 0] += 1;
    }
    for (const auto& e : edges)
      if (e.second[0] != 1 || e.second[1] != 1)
        return fail(TMPM_INVALID_INPUT, "mesh is not closed and consistently oriented");
    MeshLattice lat;
    if ((st = mesh_lattice(vertices, vertex_count, spacing, max_candidates, lat)) != TMPM_OK) return st;
    // sampling.rs:116-144 on the device: the solid-angle test of every candidate
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(TMPM_NO_DEVICE, "cannot select device");
    DevBuf dv, dt, df;
    HIP_TRY(dv.alloc(24ull * vertex_count));
    HIP_TRY(dt.alloc(12ull * triangle_count));
    HIP_TRY(df.alloc(lat.candidates));
    HIP_TRY(hipMemcpy(dv.ptr, vertices, 24ull * vertex_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dt.ptr, triangles, 12ull * triangle_count, hipMemcpyHostToDevice));
    hipStream_t stream = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    launch_mesh_inside(dv.as<double>(), dt.as<uint32_t>(), triangle_count, lat.lower, lat.upper, spacing, lat.counts,
                       df.as<uint8_t>(), stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
    if (e != hipSuccess) return fail(TMPM_READBACK, std::string("mesh sampling kernel failed: ") + hipGetErrorString(e));
    std::vector<uint8_t> flags(lat.candidates);
    HIP_TRY(hipMemcpy(flags.data(), df.ptr, lat.candidates, hipMemcpyDeviceToHost));
    uint64_t n = 0;
    if ((st = mesh_emit_samples(lat, spacing, flags.data(), 0, lat.candidates, max_samples, samples_out, n)) != TMPM_OK)
      return st;
    *count_out = n;
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

tmpm_status tmpm_sample_closed_mesh_volume_noalloc(tmpm_context* ctx, const double* vertices, uint64_t vertex_count,
                                                   const uint32_t* triangles, uint64_t triangle_count,
                                                   double spacing, uint64_t max_candidates, uint64_t max_samples,
                                                   double* samples_out, uint64_t* count_out) {
  try {
    if (count_out) *count_out = 0;
    tmpm_status st = mesh_check_triangles(ctx, vertices, vertex_count, triangles, triangle_count, spacing,
                                          max_candidates, max_samples, samples_out, count_out);
    if (st != TMPM_OK) return st;
    // scratch layout: status word, vertices, triangles, candidate flags (each 256-byte aligned)
    if (triangle_count > kNoAllocMaxTriangles)
      return fail(TMPM_CAPACITY, "mesh exceeds the allocation-free sampler's triangle limit");
    auto align = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    const uint64_t off_v = 256, off_t = align(off_v + 24 * vertex_count), off_f = align(off_t + 12 * triangle_count);
    if (vertex_count > kScratchBytes || off_f > kScratchBytes)
      return fail(TMPM_CAPACITY, "mesh exceeds the allocation-free sampler's scratch memory");
    NoAllocScratch& ns = noalloc_scratch(ctx->device);
    std::lock_guard<std::mutex> lock(ns.mu);  // calls on one device share its scratch
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(TMPM_NO_DEVICE, "cannot select device");
    if (!ns.ptr && (st = setup_scratch(ctx->device)) != TMPM_OK) return st;
    char* base = static_cast<char*>(ns.ptr);
    uint32_t* bad = reinterpret_cast<uint32_t*>(base);
    const double* dv = reinterpret_cast<const double*>(base + off_v);
    const uint32_t* dt = reinterpret_cast<const uint32_t*>(base + off_t);
    uint8_t* df = reinterpret_cast<uint8_t*>(base + off_f);
    static const uint32_t zero = 0;
    HIP_TRY(hipMemcpy(bad, &zero, sizeof(zero), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(base + off_v, vertices, 24ull * vertex_count, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(base + off_t, triangles, 12ull * triangle_count, hipMemcpyHostToDevice));
    // sampling.rs:66-93 on the device (the allocating entry point builds a host edge map)
    launch_mesh_edges(dt, triangle_count, bad, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(TMPM_READBACK, std::string("mesh edge kernel failed: ") + hipGetErrorString(e));
    uint32_t bad_host = 0;
    HIP_TRY(hipMemcpy(&bad_host, bad, sizeof(bad_host), hipMemcpyDeviceToHost));
    if (bad_host) return fail(TMPM_INVALID_INPUT, "mesh is not closed and consistently oriented");
    MeshLattice lat;
    if ((st = mesh_lattice(vertices, vertex_count, spacing, max_candidates, lat)) != TMPM_OK) return st;
    if (lat.candidates > kScratchBytes - off_f)
      return fail(TMPM_CAPACITY, "candidate lattice exceeds the allocation-free sampler's scratch memory");
    launch_mesh_inside(dv, dt, triangle_count, lat.lower, lat.upper, spacing, lat.counts, df, nullptr);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(TMPM_READBACK, std::string("mesh sampling kernel failed: ") + hipGetErrorString(e));
    // the flags come back through a fixed host buffer, chunk by chunk
    uint8_t chunk[1 << 16];
    uint64_t n = 0;
    for (uint64_t c0 = 0; c0 < lat.candidates; c0 += sizeof(chunk)) {
      const uint64_t len = std::min<uint64_t>(sizeof(chunk), lat.candidates - c0);
      HIP_TRY(hipMemcpy(chunk, df + c0, len, hipMemcpyDeviceToHost));
      if ((st = mesh_emit_samples(lat, spacing, chunk, c0, len, max_samples, samples_out, n)) != TMPM_OK) return st;
    }
    *count_out = n;
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

tmpm_status tmpm_bandwidth_probe(int32_t device_ordinal, uint64_t bytes, uint32_t iterations, double* gbs_out) {
  try {
    if (!gbs_out || bytes < 16 || iterations == 0) return fail(TMPM_INVALID_INPUT, "invalid probe arguments");
    if (hipSetDevice(device_ordinal) != hipSuccess) return fail(TMPM_NO_DEVICE, "cannot select device");
    size_t n = bytes / 16;
    DevBuf a, b;
    HIP_TRY(a.alloc(n * 16));
    HIP_TRY(b.alloc(n * 16));
    HIP_TRY(hipMemset(a.ptr, 0, n * 16));
    hipStream_t st;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    launch_copy(a.as<float4>(), b.as<float4>(), n, st);  // warm up
    HIP_TRY(hipEventRecord(e0, st));
    for (uint32_t i = 0; i < iterations; ++i) launch_copy(a.as<float4>(), b.as<float4>(), n, st);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    *gbs_out = 2.0 * (double)n * 16.0 * iterations / (ms * 1e-3) / 1e9;
    return TMPM_OK;
  } catch (...) {
    return exception_status();
  }
}

double tmpm_timestep_bound(const tmpm_material* material, double density, const double velocity[3], double cell_width,
                           double deformation_det) {
  return timestep_bound(*material, density, velocity, cell_width, deformation_det);
}
int32_t tmpm_material_valid(const tmpm_material* material) { return material && material_valid(*material) ? 1 : 0; }
int32_t tmpm_obstacle_valid(const tmpm_obstacle* obstacle) { return obstacle && obstacle_valid(*obstacle) ? 1 : 0; }

}  // extern "C"
