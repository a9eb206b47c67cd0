// TEST INFRASTRUCTURE ONLY — the FP64 oracle for the MLS-MPM substep.
//
// This file restates, in plain C++ on the CPU, the reference CPU pipeline of
// tessera-mpm (crates/tessera-mpm/src/world.rs, material.rs, obstacle.rs) and
// the f32 cell-key rule its GPU backend uses (gpu.rs:1142-1154,
// gpu_topology.wgsl:18-66).  It is the checker for the HIP backend: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark may load
// it.  Nothing in the product library links or calls it.
//
// Parity pinning (what fixes each function independently of this text):
//   * substep/advance/stable_timestep: closed forms restated from the
//     reference tests (free fall, uniform-velocity APIC, transfer colours,
//     CPIC sides, reaction momentum balance, fixed particles, walls, CFL
//     monotonicity) plus invariants (partition of unity, mass and momentum
//     conservation, affine-field reproduction) in tests/test_oracle_*.py.
//   * materials: closed forms from material.rs / sand_neo_hookean_tests.rs
//     (sand hardening sequence, apex projection, snow clamp, fluid signs,
//     isotropic roundoff) and uniqueness of the polar factor for nonsingular F
//     (R orthogonal, R^T F symmetric positive definite).
//   * obstacles: the analytic SDF values of obstacle.rs:794-1061.
//   * binning: brute-force set construction on tiny inputs and the exact
//     live-node counts of gpu.rs:3183-3186 (63, 81, 0).
//   * nalgebra's SVD for singular or repeated-singular-value F is not unique;
//     there the oracle's own one-sided Jacobi SVD is used and parity is
//     "unpinned" (documented beside the tests that touch it).
//
// Build: oracle/build_oracle.sh (g++ -O2 -ffp-contract=off, no fast-math, so
// f32 and f64 arithmetic round like the reference's Rust code).

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Small fixed-size linear algebra (nalgebra semantics: column vectors,
// Matrix3 indexed (row, column); products accumulate k = 0, 1, 2 in order).
// ---------------------------------------------------------------------------
struct V3 {
  double x = 0, y = 0, z = 0;
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double norm_squared(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(norm_squared(a)); }
inline V3 normalize(V3 a) { return a / norm(a); }
inline bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

struct M3 {
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double& operator()(int r, int c) { return a[r][c]; }
  double operator()(int r, int c) const { return a[r][c]; }
};
inline M3 identity() { M3 m; m.a[0][0] = m.a[1][1] = m.a[2][2] = 1.0; return m; }
inline M3 operator+(const M3& x, const M3& y) {
  M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] + y.a[i][j]; return r;
}
inline M3 operator-(const M3& x, const M3& y) {
  M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] - y.a[i][j]; return r;
}
inline M3 operator*(const M3& x, double s) {
  M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] * s; return r;
}
inline M3 operator*(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = x.a[i][0] * y.a[0][j];
      s += x.a[i][1] * y.a[1][j];
      s += x.a[i][2] * y.a[2][j];
      r.a[i][j] = s;
    }
  return r;
}
inline V3 operator*(const M3& m, V3 v) {
  V3 r;
  for (int i = 0; i < 3; ++i) {
    double s = m.a[i][0] * v.x;
    s += m.a[i][1] * v.y;
    s += m.a[i][2] * v.z;
    r[i] = s;
  }
  return r;
}
inline M3 transpose(const M3& m) {
  M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[j][i]; return r;
}
inline M3 outer(V3 u, V3 v) {
  M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = u[i] * v[j]; return r;
}
inline M3 diag(V3 d) { M3 m; m.a[0][0] = d.x; m.a[1][1] = d.y; m.a[2][2] = d.z; return m; }
inline double trace(const M3& m) { return m.a[0][0] + m.a[1][1] + m.a[2][2]; }
inline bool finite_m(const M3& m) {
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) if (!std::isfinite(m.a[i][j])) return false;
  return true;
}
// nalgebra's closed-form 3x3 determinant (cofactor expansion along row 0).
inline double determinant(const M3& m) {
  double m11 = m.a[0][0], m12 = m.a[0][1], m13 = m.a[0][2];
  double m21 = m.a[1][0], m22 = m.a[1][1], m23 = m.a[1][2];
  double m31 = m.a[2][0], m32 = m.a[2][1], m33 = m.a[2][2];
  double minor_m12_m23 = m22 * m33 - m32 * m23;
  double minor_m11_m23 = m21 * m33 - m31 * m23;
  double minor_m11_m22 = m21 * m32 - m31 * m22;
  return m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
}
inline double frobenius(const M3& m) {
  double s = 0; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += m.a[i][j] * m.a[i][j];
  return std::sqrt(s);
}
// Column-major 9-element storage (nalgebra's memory order) <-> M3.
inline M3 load_cm(const double* p) {
  M3 m; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) m.a[r][c] = p[c * 3 + r]; return m;
}
inline void store_cm(const M3& m, double* p) {
  for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) p[c * 3 + r] = m.a[r][c];
}

// One-sided (Hestenes) Jacobi SVD in f64: F = U diag(s) V^T with s >= 0.
// Stand-in for nalgebra 0.30.1 `Matrix3::svd(true, true)` (material.rs:534).
// For nonsingular F with distinct singular values every quantity the
// reference derives from the SVD (U V^T, U f(s) V^T) is unique, so parity
// there is pinned by uniqueness; for degenerate F it is unpinned.
void svd3(const M3& F, M3& U, V3& s, M3& V) {
  M3 A = F;
  V = identity();
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += A.a[i][p] * A.a[i][p];
          beta += A.a[i][q] * A.a[i][q];
          gamma += A.a[i][p] * A.a[i][q];
        }
        if (gamma == 0.0) continue;
        double scale = std::sqrt(alpha * beta);
        if (scale > 0) off = std::max(off, std::fabs(gamma) / scale);
        if (std::fabs(gamma) <= 1e-300) continue;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t);
        double sn = c * t;
        for (int i = 0; i < 3; ++i) {
          double ap = A.a[i][p], aq = A.a[i][q];
          A.a[i][p] = c * ap - sn * aq;
          A.a[i][q] = sn * ap + c * aq;
          double vp = V.a[i][p], vq = V.a[i][q];
          V.a[i][p] = c * vp - sn * vq;
          V.a[i][q] = sn * vp + c * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double smax = 0;
  for (int j = 0; j < 3; ++j) {
    double n2 = 0;
    for (int i = 0; i < 3; ++i) n2 += A.a[i][j] * A.a[i][j];
    s[j] = std::sqrt(n2);
    smax = std::max(smax, s[j]);
  }
  // Left singular vectors; complete the basis for (numerically) null columns.
  bool have[3] = {false, false, false};
  for (int j = 0; j < 3; ++j) {
    if (s[j] > smax * 1e-14 && s[j] > 0) {
      for (int i = 0; i < 3; ++i) U.a[i][j] = A.a[i][j] / s[j];
      have[j] = true;
    }
  }
  for (int j = 0; j < 3; ++j) {
    if (have[j]) continue;
    V3 cand = v3(0, 0, 0);
    int others[2], k = 0;
    for (int o = 0; o < 3; ++o) if (o != j) others[k++] = o;
    if (have[others[0]] && have[others[1]]) {
      V3 a = v3(U.a[0][others[0]], U.a[1][others[0]], U.a[2][others[0]]);
      V3 b = v3(U.a[0][others[1]], U.a[1][others[1]], U.a[2][others[1]]);
      cand = cross(a, b);
      // Orientation so U diag(s) V^T reproduces F is irrelevant for s_j = 0.
    } else {
      for (int e = 0; e < 3 && norm(cand) < 0.5; ++e) {
        V3 t = v3(e == 0, e == 1, e == 2);
        for (int o = 0; o < 3; ++o)
          if (have[o]) {
            V3 u = v3(U.a[0][o], U.a[1][o], U.a[2][o]);
            t = t - u * dot(u, t);
          }
        if (norm(t) > 0.5) cand = t;
      }
    }
    cand = normalize(cand);
    for (int i = 0; i < 3; ++i) U.a[i][j] = cand[i];
    have[j] = true;
  }
}

// ---------------------------------------------------------------------------
// Materials (material.rs)
// ---------------------------------------------------------------------------
enum MatKind { LINEAR_ELASTIC = 0, NEO_HOOKEAN = 1, SAND = 2, FLUID = 3, SNOW = 4, SAND_NEO_HOOKEAN = 5 };

struct Material {
  int kind = LINEAR_ELASTIC;
  double p[5] = {1000.0, 0.2, 0, 0, 0};
};
struct Plastic {
  double plastic_det = 1.0, hardening = 0.0, log_volume_gain = 0.0;
};

// material.rs:429-431
bool valid_elasticity(double young, double poisson) {
  return std::isfinite(young) && young > 0.0 && std::isfinite(poisson) && poisson >= -1.0 && poisson < 0.5;
}
// material.rs:172-231
bool material_valid(const Material& m) {
  const double* p = m.p;
  switch (m.kind) {
    case LINEAR_ELASTIC:
    case NEO_HOOKEAN: return valid_elasticity(p[0], p[1]);
    case SAND:
    case SAND_NEO_HOOKEAN:
      return valid_elasticity(p[0], p[1]) && std::isfinite(p[2]) && p[2] >= 0.0 && p[2] < M_PI / 2 &&
             std::isfinite(p[3]) && p[3] >= 0.0;
    case FLUID:
      return std::isfinite(p[0]) && p[0] > 0 && std::isfinite(p[1]) && p[1] > 0 && std::isfinite(p[2]) &&
             p[2] >= 0 && std::isfinite(p[3]) && p[3] >= 0;
    case SNOW:
      return valid_elasticity(p[0], p[1]) && std::isfinite(p[2]) && p[2] >= 0 && p[2] < 1.0 &&
             std::isfinite(p[3]) && p[3] >= 0 && std::isfinite(p[4]) && p[4] >= 0;
  }
  return false;
}
// material.rs:433-438
void lame(double young, double poisson, double& lambda, double& mu) {
  lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  mu = young / (2.0 * (1.0 + poisson));
}
// material.rs:440-455; R = U V^T from the SVD.
M3 corotated_stress(const M3& F, double young, double poisson, double factor) {
  double lambda, mu;
  lame(young, poisson, lambda, mu);
  M3 U, V;
  V3 s;
  svd3(F, U, s, V);
  M3 rotation = U * transpose(V);
  double j = determinant(F);
  M3 stress = ((F - rotation) * transpose(F)) * (2.0 * mu * factor);
  stress = stress + identity() * (lambda * factor * (j - 1.0) * j);
  return stress;
}
// material.rs:457-467
M3 neo_hookean_stress(const M3& F, double young, double poisson, double factor) {
  double lambda, mu;
  lame(young, poisson, lambda, mu);
  double j = std::max(determinant(F), 1e-10);
  return (F * transpose(F)) * (mu * factor) + identity() * (factor * (lambda * std::log(j) - mu));
}
// material.rs:469-486
M3 fluid_stress(const M3& F, const M3& L, double bulk, double gamma, double viscosity, double tensile) {
  double j = std::max(determinant(F), 1e-6);
  double pressure = j <= 1.0 ? bulk * (std::pow(j, -gamma) - 1.0) : -bulk * tensile * (j - 1.0);
  M3 strain_rate = (L + transpose(L)) * 0.5;
  M3 deviatoric = strain_rate - identity() * (trace(strain_rate) / 3.0);
  return deviatoric * (2.0 * viscosity * j) - identity() * (pressure * j);
}
// material.rs:234-284
M3 kirchhoff_stress(const Material& m, const M3& F, const M3& L, const Plastic& pl) {
  const double* p = m.p;
  switch (m.kind) {
    case LINEAR_ELASTIC: return corotated_stress(F, p[0], p[1], 1.0);
    case NEO_HOOKEAN:
    case SAND_NEO_HOOKEAN: return neo_hookean_stress(F, p[0], p[1], 1.0);
    case SAND: return corotated_stress(F, p[0], p[1], 1.0);
    case FLUID: return fluid_stress(F, L, p[0], p[1], p[2], p[3]);
    case SNOW: {
      double factor = std::exp(p[4] * (1.0 - pl.plastic_det));
      factor = std::min(std::max(factor, 0.01), 100.0);
      return corotated_stress(F, p[0], p[1], factor);
    }
  }
  return M3();
}
inline V3 vmap_exp(V3 v) { return v3(std::exp(v.x), std::exp(v.y), std::exp(v.z)); }
// material.rs:488-532
void project_sand(M3& F, Plastic& pl, double young, double poisson, double friction_angle, double cohesion) {
  double lambda, mu;
  lame(young, poisson, lambda, mu);
  if (lambda == 0.0) return;
  M3 U, V;
  V3 s;
  svd3(F, U, s, V);
  V3 strain = v3(std::log(std::max(s.x, 1e-12)), std::log(std::max(s.y, 1e-12)), std::log(std::max(s.z, 1e-12)));
  strain = strain + v3(1, 1, 1) * (pl.log_volume_gain / 3.0);
  double tr = strain.x + strain.y + strain.z;
  V3 deviatoric = strain - v3(1, 1, 1) * (tr / 3.0);
  double nrm = norm(deviatoric);
  double shifted_trace = tr - cohesion;
  double q = pl.hardening;
  double deg = M_PI / 180.0;
  double angle = friction_angle + (9.0 * deg * q - 10.0 * deg) * std::exp(-0.2 * q);
  double sine = std::sin(angle);
  double alpha = std::sqrt(2.0 / 3.0) * (2.0 * sine) / (3.0 - sine);
  V3 projected_log;
  double increment;
  if (shifted_trace > 0.0 || nrm <= 4.0 * (double)std::numeric_limits<float>::epsilon()) {
    projected_log = v3(1, 1, 1) * (cohesion / 3.0);
    increment = norm(strain);
  } else {
    double gamma = nrm + (3.0 * lambda + 2.0 * mu) / (2.0 * mu) * shifted_trace * alpha;
    if (gamma <= 0.0) return;
    projected_log = strain - deviatoric * (gamma / nrm);
    increment = gamma;
  }
  V3 projected = vmap_exp(projected_log);
  double old_det = std::max(s.x * s.y * s.z, 1e-12);
  double new_det = std::max(projected.x * projected.y * projected.z, 1e-12);
  pl.plastic_det = std::min(std::max(pl.plastic_det * old_det / new_det, 0.01), 100.0);
  pl.log_volume_gain += std::log(old_det) - std::log(new_det);
  pl.hardening += increment;
  F = U * diag(projected) * transpose(V);
}
// material.rs:287-371
void project_deformation(const Material& m, M3& F, Plastic& pl) {
  const double* p = m.p;
  switch (m.kind) {
    case FLUID: {
      double j = std::max(determinant(F), 1e-9);
      F = identity() * std::cbrt(j);
      return;
    }
    case SNOW: {
      M3 U, V;
      V3 s;
      svd3(F, U, s, V);
      V3 c;
      for (int i = 0; i < 3; ++i) c[i] = std::min(std::max(s[i], 1.0 - p[2]), 1.0 + p[3]);
      double old_det = s.x * s.y * s.z;
      double new_det = std::max(c.x * c.y * c.z, 1e-12);
      pl.plastic_det = std::min(std::max(pl.plastic_det * old_det / new_det, 0.1), 4.0);
      F = U * diag(c) * transpose(V);
      return;
    }
    case SAND: project_sand(F, pl, p[0], p[1], p[2], p[3]); return;
    case SAND_NEO_HOOKEAN: {
      double lambda, mu;
      lame(p[0], p[1], lambda, mu);
      if (lambda == 0.0) return;
      M3 U, V;
      V3 s;
      svd3(F, U, s, V);
      V3 strain = v3(std::log(s.x), std::log(s.y), std::log(s.z)) + v3(1, 1, 1) * (pl.log_volume_gain / 3.0);
      double tr = strain.x + strain.y + strain.z;
      V3 deviatoric = strain - v3(1, 1, 1) * (tr / 3.0);
      double nrm = norm(deviatoric);
      double shifted_trace = tr - p[3];
      double q = pl.hardening;
      double deg = M_PI / 180.0;
      double angle = p[2] + (9.0 * deg * q - 10.0 * deg) * std::exp(-0.2 * q);
      double sine = std::sin(angle);
      double alpha = std::sqrt(2.0 / 3.0) * (2.0 * sine) / (3.0 - sine);
      bool isotropic = nrm <= 4.0 * (double)std::numeric_limits<float>::epsilon();
      V3 projected_log;
      double increment;
      if (shifted_trace > 0.0 || isotropic) {
        projected_log = v3(1, 1, 1) * (p[3] / 3.0);
        increment = norm(strain);
      } else {
        double gamma = nrm + (3.0 * lambda + 2.0 * mu) / (2.0 * mu) * shifted_trace * alpha;
        if (gamma <= 0.0) return;
        projected_log = strain - deviatoric * (gamma / nrm);
        increment = gamma;
      }
      V3 projected = vmap_exp(projected_log);
      double old_det = s.x * s.y * s.z;
      double new_det = projected.x * projected.y * projected.z;
      pl.plastic_det *= old_det / new_det;
      pl.log_volume_gain += std::log(old_det) - std::log(new_det);
      pl.hardening += increment;
      F = U * diag(projected) * transpose(V);
      return;
    }
    default: return;  // LinearElastic, NeoHookean: identity projection
  }
}
// material.rs:380-426
double timestep_bound(const Material& m, double density, V3 velocity, double h, double det) {
  const double* p = m.p;
  double wave_speed;
  double lambda, mu;
  switch (m.kind) {
    case SAND_NEO_HOOKEAN: {
      lame(p[0], p[1], lambda, mu);
      double d = density / std::max(det, 1e-6);
      double ws = std::sqrt((lambda + 2.0 * mu) / d);
      return 0.5 * h / std::max(ws, norm(velocity));
    }
    case FLUID: wave_speed = std::sqrt(p[0] * p[1] / std::max(density, 1e-12)); break;
    default:
      lame(p[0], p[1], lambda, mu);
      wave_speed = std::sqrt((lambda + 2.0 * mu) / std::max(density, 1e-12));
  }
  return 0.4 * h / std::max(wave_speed + norm(velocity), 1e-12);
}

// ---------------------------------------------------------------------------
// Obstacles (obstacle.rs)
// ---------------------------------------------------------------------------
enum Shape { SPHERE = 0, BOX = 1, CAPSULE = 2, CYLINDER = 3, CONE = 4, GROUND = 5, TRIANGLE_PRISM = 6, CONVEX = 7 };
enum Boundary { SLIP = 0, STICK = 1, SEPARATE = 2, NON_REFLECTING = 3, FREE_SLIP = 4 };

struct Quat { double i = 0, j = 0, k = 0, w = 1; };
// nalgebra UnitQuaternion * Vector3: t = 2 (q.v x v); v + w t + q.v x t.
inline V3 rotate(const Quat& q, V3 v) {
  V3 qv = v3(q.i, q.j, q.k);
  V3 t = cross(qv, v) * 2.0;
  V3 c = cross(qv, t);
  return t * q.w + c + v;
}
inline V3 inverse_rotate(const Quat& q, V3 v) {
  Quat c; c.i = -q.i; c.j = -q.j; c.k = -q.k; c.w = q.w;
  return rotate(c, v);
}

struct Obstacle {
  int shape = SPHERE, boundary = SLIP, cpic_group = -1;
  V3 center;
  Quat orientation;
  V3 linear_velocity, angular_velocity;
  double friction = 0.0;
  double dims[4] = {0, 0, 0, 0};
  V3 vertices[3];
  std::vector<double> planes;  // 4 per plane
  double bound_radius = 0.0;
};

// obstacle.rs:265-341
bool obstacle_valid(const Obstacle& o) {
  bool shape_valid = false;
  const double* d = o.dims;
  switch (o.shape) {
    case SPHERE: shape_valid = std::isfinite(d[0]) && d[0] > 0; break;
    case BOX: shape_valid = std::isfinite(d[0]) && d[0] > 0 && std::isfinite(d[1]) && d[1] > 0 && std::isfinite(d[2]) && d[2] > 0; break;
    case CAPSULE:
      shape_valid = std::isfinite(d[0]) && d[0] >= 0 && std::isfinite(d[1]) && d[1] >= 0 && (d[0] > 0 || d[1] > 0);
      break;
    case CYLINDER:
    case CONE: shape_valid = std::isfinite(d[0]) && d[0] > 0 && std::isfinite(d[1]) && d[1] > 0; break;
    case GROUND: shape_valid = std::isfinite(d[0]) && d[0] > 0 && std::isfinite(d[1]) && d[1] > 0; break;
    case TRIANGLE_PRISM: {
      double area2 = norm_squared(cross(o.vertices[1] - o.vertices[0], o.vertices[2] - o.vertices[0]));
      shape_valid = std::isfinite(d[0]) && d[0] > 0 && finite3(o.vertices[0]) && finite3(o.vertices[1]) &&
                    finite3(o.vertices[2]) && std::isfinite(area2) && area2 > 1e-24;
      break;
    }
    case CONVEX: {
      size_t n = o.planes.size() / 4;
      shape_valid = n >= 4 && std::isfinite(o.bound_radius) && o.bound_radius > 0;
      for (size_t i = 0; i < n && shape_valid; ++i) {
        const double* pl = &o.planes[4 * i];
        for (int k = 0; k < 4; ++k) shape_valid = shape_valid && std::isfinite(pl[k]);
        double n2 = std::fma(pl[0], pl[0], std::fma(pl[1], pl[1], pl[2] * pl[2]));
        shape_valid = shape_valid && std::fabs(n2 - 1.0) <= 1e-4;
      }
      break;
    }
    default: return false;
  }
  bool cpic_ok = o.cpic_group < 0 || (o.cpic_group < 32 && o.shape == TRIANGLE_PRISM);
  return shape_valid && cpic_ok && finite3(o.center) && std::isfinite(o.orientation.i) &&
         std::isfinite(o.orientation.j) && std::isfinite(o.orientation.k) && std::isfinite(o.orientation.w) &&
         finite3(o.linear_velocity) && finite3(o.angular_velocity) && std::isfinite(o.friction) && o.friction >= 0;
}

// obstacle.rs:344-370
bool cpic_side(const Obstacle& o, V3 point, double& dist, bool& positive) {
  if (o.shape != TRIANGLE_PRISM) return false;
  V3 local = inverse_rotate(o.orientation, point - o.center);
  V3 a = o.vertices[0];
  V3 eb = o.vertices[1] - a, ec = o.vertices[2] - a;
  V3 n = normalize(cross(eb, ec));
  double signed_d = dot(local - a, n);
  V3 projected = local - n * signed_d - a;
  double d00 = dot(eb, eb), d01 = dot(eb, ec), d11 = dot(ec, ec);
  double d20 = dot(projected, eb), d21 = dot(projected, ec);
  double den = d00 * d11 - d01 * d01;
  double v = (d11 * d20 - d01 * d21) / den;
  double w = (d00 * d21 - d01 * d20) / den;
  if (v >= -1e-8 && w >= -1e-8 && v + w <= 1.0 + 1e-8) {
    dist = std::fabs(signed_d);
    positive = signed_d >= 0.0;
    return true;
  }
  return false;
}

// obstacle.rs:542-586
void cylinder_surface(V3 local, double hh, double r, double& dist, V3& normal) {
  double radial = std::sqrt(local.x * local.x + local.y * local.y);
  V3 radial_normal = radial > 1e-12 ? v3(local.x / radial, local.y / radial, 0.0) : v3(1, 0, 0);
  double radial_distance = radial - r;
  double cap_distance = std::fabs(local.z) - hh;
  double eps = std::max(hh, r) * 1e-6;
  if (std::fabs(radial_distance) <= eps && std::fabs(cap_distance) <= eps) {
    V3 cap_normal = local.z < 0.0 ? v3(0, 0, -1) : v3(0, 0, 1);
    dist = std::max(radial_distance, cap_distance);
    normal = normalize(radial_normal + cap_normal);
    return;
  }
  double ro = std::max(radial_distance, 0.0);
  double co = std::max(cap_distance, 0.0);
  V3 outside = radial_normal * ro + v3(0, 0, 1) * (local.z < 0.0 ? -co : co);
  double ol = norm(outside);
  if (ol > 1e-12) {
    dist = ol;
    normal = outside / ol;
  } else if (radial_distance >= cap_distance) {
    dist = radial_distance;
    normal = radial_normal;
  } else {
    dist = cap_distance;
    normal = local.z < 0.0 ? v3(0, 0, -1) : v3(0, 0, 1);
  }
}

// obstacle.rs:589-630
void cone_surface(V3 local, double hh, double r, double& dist, V3& normal) {
  double radial = std::sqrt(local.x * local.x + local.y * local.y);
  V3 radial_normal = radial > 1e-12 ? v3(local.x / radial, local.y / radial, 0.0) : v3(1, 0, 0);
  double px = radial, pz = local.z;
  double bx = std::min(std::max(radial, 0.0), r), bz = -hh;
  double sx = r, sz = -hh;
  double dx = -r, dz = 2.0 * hh;
  double t = ((px - sx) * dx + (pz - sz) * dz) / (dx * dx + dz * dz);
  t = std::min(std::max(t, 0.0), 1.0);
  double cx = sx + dx * t, cz = sz + dz * t;
  double eps = std::max(hh, r) * 1e-6;
  double bd2 = (px - bx) * (px - bx) + (pz - bz) * (pz - bz);
  double sd2 = (px - cx) * (px - cx) + (pz - cz) * (pz - cz);
  bool use_base = bd2 <= sd2;
  double clx = use_base ? bx : cx, clz = use_base ? bz : cz;
  double ddx = px - clx, ddz = pz - clz;
  double nrm = std::sqrt(ddx * ddx + ddz * ddz);
  bool inside = local.z >= -hh - eps && local.z <= hh + eps && radial <= r * (hh - local.z) / (2.0 * hh) + eps;
  double bnx = 0.0, bnz = -1.0;
  double snl = std::sqrt(4.0 * hh * hh + r * r);
  double snx = 2.0 * hh / snl, snz = r / snl;
  bool near_corner = bd2 <= eps * eps && sd2 <= eps * eps;
  double fnx, fnz;
  if (near_corner) {
    double ax = bnx + snx, az = bnz + snz;
    double l = std::sqrt(ax * ax + az * az);
    fnx = ax / l;
    fnz = az / l;
  } else if (use_base) {
    fnx = bnx; fnz = bnz;
  } else {
    fnx = snx; fnz = snz;
  }
  double n2x, n2z;
  if (inside || nrm <= eps) { n2x = fnx; n2z = fnz; } else { n2x = ddx / nrm; n2z = ddz / nrm; }
  normal = radial_normal * n2x + v3(0, 0, 1) * n2z;
  dist = inside ? -nrm : nrm;
}

// obstacle.rs:747-788
V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  V3 bp = p - b;
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + ab * (d1 / (d1 - d3));
  V3 cp = p - c;
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + ac * (d2 / (d2 - d6));
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  double rec = 1.0 / (va + vb + vc);
  return a + ab * (vb * rec) + ac * (vc * rec);
}

// obstacle.rs:633-694
void triangle_prism_surface(V3 point, const V3* v, double ht, double& dist, V3& outward) {
  V3 a = v[0], b = v[1], c = v[2];
  V3 n = normalize(cross(b - a, c - a));
  V3 off = n * ht;
  V3 prism[6] = {a + off, b + off, c + off, a - off, b - off, c - off};
  static const int faces[8][3] = {{0, 1, 2}, {3, 5, 4}, {3, 4, 1}, {3, 1, 0}, {4, 5, 2}, {4, 2, 1}, {5, 3, 0}, {5, 0, 2}};
  double edge_scale = std::max(std::max(std::max(norm(b - a), norm(c - b)), norm(a - c)), ht);
  double eps = edge_scale * 1e-6;
  bool inside = true;
  V3 closest;
  double best = std::numeric_limits<double>::infinity();
  V3 best_normal = n, tied = v3(0, 0, 0);
  for (auto& f : faces) {
    V3 fn = normalize(cross(prism[f[1]] - prism[f[0]], prism[f[2]] - prism[f[0]]));
    inside = inside && (dot(point - prism[f[0]], fn) <= eps);
    V3 cand = closest_point_triangle(point, prism[f[0]], prism[f[1]], prism[f[2]]);
    double d2 = norm_squared(point - cand);
    if (d2 + eps * eps < best) {
      best = d2;
      closest = cand;
      best_normal = fn;
      tied = fn;
    } else if (std::fabs(d2 - best) <= eps * eps) {
      tied = tied + fn;
    }
  }
  double distance = std::sqrt(best);
  double tl = norm(tied);
  V3 feature = tl > 1e-12 ? tied / tl : best_normal;
  outward = (inside || distance <= eps) ? feature : (point - closest) / distance;
  dist = inside ? -distance : distance;
}

// obstacle.rs:696-745
void convex_surface(V3 point, const std::vector<double>& planes, double bound_radius, double& dist, V3& normal) {
  size_t n = planes.size() / 4;
  double maximum = -std::numeric_limits<double>::infinity();
  V3 nearest = v3(0, 0, 1);
  for (size_t i = 0; i < n; ++i) {
    V3 nn = v3(planes[4 * i], planes[4 * i + 1], planes[4 * i + 2]);
    double d = dot(nn, point) - planes[4 * i + 3];
    if (d > maximum) { maximum = d; nearest = nn; }
  }
  if (maximum <= 0.0) { dist = maximum; normal = nearest; return; }
  V3 projected = point;
  std::vector<V3> corr(n, v3(0, 0, 0));
  double tol = std::max(bound_radius, 1.0) * 1e-10;
  for (int it = 0; it < 128; ++it) {
    V3 prev = projected;
    for (size_t i = 0; i < n; ++i) {
      V3 nn = v3(planes[4 * i], planes[4 * i + 1], planes[4 * i + 2]);
      V3 shifted = projected + corr[i];
      double viol = std::max(dot(nn, shifted) - planes[4 * i + 3], 0.0);
      projected = shifted - nn * viol;
      corr[i] = shifted - projected;
    }
    if (norm(projected - prev) <= tol) break;
  }
  V3 off = point - projected;
  double d = norm(off);
  if (d > tol) { dist = d; normal = off / d; } else { dist = maximum; normal = nearest; }
}

// obstacle.rs:373-472
void surface(const Obstacle& o, V3 point, double& dist, V3& normal) {
  V3 rel = point - o.center;
  const double* d = o.dims;
  switch (o.shape) {
    case SPHERE: {
      double nr = norm(rel);
      normal = nr > 1e-12 ? rel / nr : v3(0, 0, 1);
      dist = nr - d[0];
      return;
    }
    case BOX: {
      V3 local = inverse_rotate(o.orientation, rel);
      V3 cl = v3(std::min(std::max(local.x, -d[0]), d[0]), std::min(std::max(local.y, -d[1]), d[1]),
                 std::min(std::max(local.z, -d[2]), d[2]));
      V3 outside = local - cl;
      double nr = norm(outside);
      if (nr > 1e-12) {
        dist = nr;
        normal = rotate(o.orientation, outside / nr);
      } else {
        V3 dd = v3(d[0] - std::fabs(local.x), d[1] - std::fabs(local.y), d[2] - std::fabs(local.z));
        int axis = (dd.x <= dd.y && dd.x <= dd.z) ? 0 : (dd.y <= dd.z ? 1 : 2);
        V3 nn = v3(0, 0, 0);
        nn[axis] = local[axis] < 0.0 ? -1.0 : 1.0;
        dist = -dd[axis];
        normal = rotate(o.orientation, nn);
      }
      return;
    }
    case CAPSULE: {
      V3 local = inverse_rotate(o.orientation, rel);
      double cz = std::min(std::max(local.z, -d[0]), d[0]);
      V3 delta = local - v3(0, 0, cz);
      double nr = norm(delta);
      double eps = std::max(d[0], d[1]) * 1e-6;
      V3 nn = nr > eps ? delta / nr : v3(1, 0, 0);
      dist = nr - d[1];
      normal = rotate(o.orientation, nn);
      return;
    }
    case CYLINDER: {
      V3 local = inverse_rotate(o.orientation, rel), nn;
      cylinder_surface(local, d[0], d[1], dist, nn);
      normal = rotate(o.orientation, nn);
      return;
    }
    case CONE: {
      V3 local = inverse_rotate(o.orientation, rel), nn;
      cone_surface(local, d[0], d[1], dist, nn);
      normal = rotate(o.orientation, nn);
      return;
    }
    case GROUND: {
      V3 local = inverse_rotate(o.orientation, rel);
      double ox = local.x - std::min(std::max(local.x, -d[0]), d[0]);
      double oy = local.y - std::min(std::max(local.y, -d[1]), d[1]);
      V3 outside = v3(ox, oy, local.z);
      double horizontal = std::hypot(ox, oy);
      if (horizontal <= 1e-12) {
        dist = local.z;
        normal = rotate(o.orientation, v3(0, 0, 1));
      } else {
        double dd = norm(outside);
        dist = dd;
        normal = rotate(o.orientation, outside / dd);
      }
      return;
    }
    case TRIANGLE_PRISM: {
      V3 local = inverse_rotate(o.orientation, rel), nn;
      triangle_prism_surface(local, o.vertices, d[0], dist, nn);
      normal = rotate(o.orientation, nn);
      return;
    }
    case CONVEX: {
      V3 local = inverse_rotate(o.orientation, rel), nn;
      convex_surface(local, o.planes, o.bound_radius, dist, nn);
      normal = rotate(o.orientation, nn);
      return;
    }
  }
}

// obstacle.rs:475-499
bool may_contact(const Obstacle& o, V3 point, double margin) {
  switch (o.shape) {
    case CAPSULE: {
      double b = o.dims[0] + o.dims[1] + margin;
      return norm_squared(point - o.center) <= b * b;
    }
    case TRIANGLE_PRISM: {
      double r = std::max(std::max(std::max(0.0, norm(o.vertices[0])), norm(o.vertices[1])), norm(o.vertices[2])) +
                 o.dims[0] + margin;
      return norm_squared(point - o.center) <= r * r;
    }
    case CONVEX: {
      double b = o.bound_radius + margin;
      return norm_squared(point - o.center) <= b * b;
    }
    default: return true;
  }
}

// obstacle.rs:502-539
V3 contact_velocity(const Obstacle& o, V3 point, V3 velocity, V3 normal) {
  V3 surface_velocity = o.linear_velocity + cross(o.angular_velocity, point - o.center);
  if (o.boundary == NON_REFLECTING) return velocity;
  if (o.boundary == STICK) return surface_velocity;
  V3 relative = velocity - surface_velocity;
  double normal_speed = dot(relative, normal);
  if (o.boundary == FREE_SLIP) return velocity - normal * normal_speed;
  if (normal_speed >= 0.0) return velocity;
  V3 tangent = relative - normal * normal_speed;
  if (o.boundary == SEPARATE) return surface_velocity + tangent;
  double ts = norm(tangent);
  double fs = ts > 1e-12 ? std::max(1.0 - o.friction * -normal_speed / ts, 0.0) : 0.0;
  return surface_velocity + tangent * fs;
}

// ---------------------------------------------------------------------------
// World (world.rs)
// ---------------------------------------------------------------------------
struct Particle {
  V3 position, velocity;
  M3 deformation, affine;
  V3 force;
  double mass = 0, rest_volume = 0, radius = 0, damping = 0;
  Material material;
  Plastic plastic;
  uint32_t group_id = 0;
  uint16_t transfer_color = 0;
  bool enabled = true, fixed = false;
  uint64_t chunk = 0;
};

struct Params {
  double cell_width = 0.1;
  V3 gravity = v3(0, 0, -9.81);
  double max_substep = 1e-3;
  bool has_bounds = false;
  V3 bmin, bmax;
};

enum Status { OK = 0, INVALID_INPUT = 1, EXCESSIVE_SUBSTEPS = 6, NON_FINITE_STATE = 7, UNKNOWN_CHUNK = 8, PARTICLE_CAPACITY = 9 };

// world.rs:53-64 (with WorldBounds::is_valid, world.rs:20-24)
bool params_valid(const Params& p) {
  if (!std::isfinite(p.cell_width) || p.cell_width <= 0.0 || !finite3(p.gravity) || !std::isfinite(p.max_substep) ||
      p.max_substep <= 0.0)
    return false;
  if (p.has_bounds) {
    if (!finite3(p.bmin) || !finite3(p.bmax)) return false;
    for (int a = 0; a < 3; ++a) if (!(p.bmax[a] - p.bmin[a] > 0.0)) return false;
  }
  return true;
}
// world.rs:505-515
bool finite_particle_state(const Particle& q) {
  return finite3(q.position) && finite3(q.velocity) && finite_m(q.deformation) && finite_m(q.affine) &&
         finite3(q.force) && std::isfinite(q.mass) && std::isfinite(q.rest_volume) && std::isfinite(q.radius) &&
         std::isfinite(q.damping);
}
// world.rs:492-503
bool valid_particle(const Particle& q) {
  return finite_particle_state(q) && q.mass > 0.0 && q.rest_volume > 0.0 && q.radius > 0.0 &&
         material_valid(q.material) && q.damping >= 0.0 && std::isfinite(q.plastic.plastic_det) &&
         q.plastic.plastic_det > 0.0 && std::isfinite(q.plastic.hardening) && std::isfinite(q.plastic.log_volume_gain);
}

struct NodeKey {
  int32_t x, y, z;
  uint64_t tk;
  bool operator==(const NodeKey& o) const { return x == o.x && y == o.y && z == o.z && tk == o.tk; }
  bool operator<(const NodeKey& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    if (z != o.z) return z < o.z;
    return tk < o.tk;
  }
};
struct NodeKeyHash {
  size_t operator()(const NodeKey& k) const {
    uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    h ^= k.tk * 0x27D4EB2F165667C5ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};
struct GridNode {
  double mass = 0;
  V3 momentum, velocity;
};
struct Reaction {
  V3 linear, angular;
};
// world.rs:200-203
inline void record(Reaction& r, V3 center, V3 point, V3 impulse) {
  r.linear = r.linear + impulse;
  r.angular = r.angular + cross(point - center, impulse);
}

// world.rs:517-523
inline void base_cell(V3 gp, int32_t b[3]) {
  b[0] = (int32_t)std::floor(gp.x - 0.5);
  b[1] = (int32_t)std::floor(gp.y - 0.5);
  b[2] = (int32_t)std::floor(gp.z - 0.5);
}
// world.rs:529-538
inline void quadratic_weights(V3 fx, double w[3][3]) {
  for (int a = 0; a < 3; ++a) {
    double v = fx[a];
    double t0 = 1.5 - v, t1 = v - 1.0, t2 = v - 0.5;
    w[a][0] = 0.5 * (t0 * t0);
    w[a][1] = 0.75 - t1 * t1;
    w[a][2] = 0.5 * (t2 * t2);
  }
}

struct World {
  std::vector<Particle> particles;
  Params params;
  std::vector<Obstacle> obstacles;
  uint64_t substeps = 0;
  bool use_cpic = true;
  uint64_t next_chunk_id = 1;

  // world.rs:470-490
  uint64_t transfer_key(const Particle& q) const {
    bool have[32] = {false};
    double best[32];
    bool side[32] = {false};
    if (use_cpic) {
      for (const Obstacle& o : obstacles) {
        if (o.cpic_group < 0) continue;
        double d;
        bool pos;
        if (cpic_side(o, q.position, d, pos)) {
          int g = o.cpic_group;
          if (!have[g] || d < best[g]) {
            have[g] = true;
            best[g] = d;
            side[g] = pos;
          }
        }
      }
    }
    uint32_t mask = 0;
    for (int g = 0; g < 32; ++g) if (have[g] && side[g]) mask |= 1u << g;
    return ((uint64_t)q.transfer_color << 32) | (uint64_t)mask;
  }

  // world.rs:289-303
  double stable_timestep() const {
    double s = params.max_substep;
    for (const Particle& q : particles) {
      if (!q.enabled || q.fixed) continue;
      double b = timestep_bound(q.material, q.mass / q.rest_volume, q.velocity, params.cell_width,
                                determinant(q.deformation));
      s = std::fmin(s, b);
    }
    return std::fmin(s, params.max_substep);
  }

  // world.rs:332-359
  int advance(double dt, std::vector<Reaction>* reactions) {
    if (!std::isfinite(dt) || dt <= 0.0) return INVALID_INPUT;
    for (const Obstacle& o : obstacles) if (!obstacle_valid(o)) return INVALID_INPUT;
    double remaining = dt;
    size_t count = 0;
    while (remaining > 0.0) {
      double stable = stable_timestep();
      if (!std::isfinite(stable) || stable <= 1e-12) return EXCESSIVE_SUBSTEPS;
      double sub = std::min(remaining, stable);
      int st = substep(sub, reactions);
      if (st != OK) return st;
      remaining = std::max(remaining - sub, 0.0);
      count += 1;
      if (count > 100000) return EXCESSIVE_SUBSTEPS;
    }
    return OK;
  }

  // world.rs:361-467
  int substep(double dt, std::vector<Reaction>* reactions) {
    const double h = params.cell_width;
    const double inv_h = 1.0 / h;
    const double inv_d = 4.0 * inv_h * inv_h;
    std::vector<uint64_t> keys(particles.size());
    for (size_t i = 0; i < particles.size(); ++i) keys[i] = transfer_key(particles[i]);
    std::unordered_map<NodeKey, uint32_t, NodeKeyHash> index;
    index.reserve(particles.size() * 8 + 16);
    std::vector<NodeKey> node_keys;
    std::vector<GridNode> nodes;

    for (size_t i = 0; i < particles.size(); ++i) {
      const Particle& q = particles[i];
      if (!q.enabled) continue;
      V3 gp = q.position * inv_h;
      int32_t b[3];
      base_cell(gp, b);
      V3 fx = gp - v3((double)b[0], (double)b[1], (double)b[2]);
      double w[3][3];
      quadratic_weights(fx, w);
      M3 stress = kirchhoff_stress(q.material, q.deformation, q.affine, q.plastic);
      M3 force_affine = stress * (-dt * q.rest_volume * inv_d);
      M3 momentum_affine = force_affine + q.affine * q.mass;
      V3 mv = q.velocity * q.mass;
      mv = v3(q.mass * q.velocity.x, q.mass * q.velocity.y, q.mass * q.velocity.z);
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c)
          for (int e = 0; e < 3; ++e) {
            NodeKey k{b[0] + a, b[1] + c, b[2] + e, keys[i]};
            double weight = w[0][a] * w[1][c] * w[2][e];
            V3 dpos = (v3((double)a, (double)c, (double)e) - fx) * h;
            auto it = index.find(k);
            uint32_t id;
            if (it == index.end()) {
              id = (uint32_t)nodes.size();
              index.emplace(k, id);
              node_keys.push_back(k);
              nodes.emplace_back();
            } else {
              id = it->second;
            }
            GridNode& n = nodes[id];
            n.mass += weight * q.mass;
            n.momentum = n.momentum + weight * (mv + momentum_affine * dpos);
          }
    }

    // Grid update in key order (BTreeMap iteration order).
    std::vector<uint32_t> order(nodes.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return node_keys[a] < node_keys[b]; });
    for (uint32_t id : order) {
      GridNode& n = nodes[id];
      if (n.mass <= 0.0) continue;
      n.velocity = n.momentum / n.mass + params.gravity * dt;
      const NodeKey& k = node_keys[id];
      int32_t cell[3] = {k.x, k.y, k.z};
      if (params.has_bounds) {
        // world.rs:561-575
        V3 pos = v3((double)cell[0], (double)cell[1], (double)cell[2]) * h;
        for (int a = 0; a < 3; ++a) {
          if ((pos[a] <= params.bmin[a] + h && n.velocity[a] < 0.0) ||
              (pos[a] >= params.bmax[a] - h && n.velocity[a] > 0.0))
            n.velocity[a] = 0.0;
        }
      }
      // world.rs:577-603
      V3 point = v3((double)cell[0], (double)cell[1], (double)cell[2]) * h;
      for (size_t oi = 0; oi < obstacles.size(); ++oi) {
        const Obstacle& o = obstacles[oi];
        if (!may_contact(o, point, h)) continue;
        double dist;
        V3 normal;
        surface(o, point, dist, normal);
        if (dist <= h * (1.0 + 1e-5)) {
          V3 before = n.velocity;
          n.velocity = contact_velocity(o, point, n.velocity, normal);
          if (reactions) record((*reactions)[oi], o.center, point - normal * dist, (before - n.velocity) * n.mass);
        }
      }
    }

    for (size_t i = 0; i < particles.size(); ++i) {
      Particle& q = particles[i];
      if (!q.enabled) {
        q.force = V3();
        continue;
      }
      if (q.fixed) {
        q.velocity = V3();
        q.affine = M3();
        q.force = V3();
        continue;
      }
      V3 gp = q.position * inv_h;
      int32_t b[3];
      base_cell(gp, b);
      V3 fx = gp - v3((double)b[0], (double)b[1], (double)b[2]);
      double w[3][3];
      quadratic_weights(fx, w);
      V3 velocity;
      M3 affine;
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c)
          for (int e = 0; e < 3; ++e) {
            NodeKey k{b[0] + a, b[1] + c, b[2] + e, keys[i]};
            double weight = w[0][a] * w[1][c] * w[2][e];
            V3 dpos = (v3((double)a, (double)c, (double)e) - fx) * h;
            auto it = index.find(k);
            V3 nv = it == index.end() ? V3() : nodes[it->second].velocity;
            velocity = velocity + nv * weight;
            affine = affine + outer(nv, dpos) * (weight * inv_d);
          }
      velocity = velocity + q.force / q.mass * dt;
      velocity = velocity * std::exp(-q.damping * dt);
      q.velocity = velocity;
      q.affine = affine;
      q.position = q.position + velocity * dt;
      q.deformation = (identity() + affine * dt) * q.deformation;
      project_deformation(q.material, q.deformation, q.plastic);
      q.force = V3();
      if (params.has_bounds) {
        // world.rs:632-644
        for (int a = 0; a < 3; ++a) {
          double lower = params.bmin[a] + q.radius;
          double upper = params.bmax[a] - q.radius;
          if (q.position[a] < lower) {
            q.position[a] = lower;
            q.velocity[a] = std::fmax(q.velocity[a], 0.0);
          } else if (q.position[a] > upper) {
            q.position[a] = upper;
            q.velocity[a] = std::fmin(q.velocity[a], 0.0);
          }
        }
      }
      // world.rs:605-630
      for (size_t oi = 0; oi < obstacles.size(); ++oi) {
        const Obstacle& o = obstacles[oi];
        if (!may_contact(o, q.position, q.radius)) continue;
        double dist;
        V3 normal;
        surface(o, q.position, dist, normal);
        if (dist < q.radius) {
          q.position = q.position + normal * (q.radius - dist);
          V3 before = q.velocity;
          q.velocity = contact_velocity(o, q.position, q.velocity, normal);
          if (reactions)
            record((*reactions)[oi], o.center, q.position - normal * q.radius, (before - q.velocity) * q.mass);
        }
      }
      if (!finite_particle_state(q)) return NON_FINITE_STATE;
    }
    substeps = substeps == UINT64_MAX ? substeps : substeps + 1;
    return OK;
  }

  // The same substep over `threads` host threads, for timing the CPU
  // baseline on all cores (no reactions).  Particles are split into
  // contiguous chunks; each thread accumulates its chunk's P2G into its own
  // grid, the grids are merged in thread order, and the grid update and G2P
  // run in parallel.  Deterministic for a given thread count; node sums are
  // associated per chunk, so the last bits differ from `substep`.
  int substep_threads(double dt, int threads) {
    const double h = params.cell_width;
    const double inv_h = 1.0 / h;
    const double inv_d = 4.0 * inv_h * inv_h;
    const size_t np = particles.size();
    const int T = std::max(1, threads);
    std::vector<uint64_t> keys(np);
    for (size_t i = 0; i < np; ++i) keys[i] = transfer_key(particles[i]);
    struct Local {
      std::unordered_map<NodeKey, uint32_t, NodeKeyHash> index;
      std::vector<NodeKey> keys;
      std::vector<GridNode> nodes;
    };
    std::vector<Local> local(T);
    auto chunk = [&](int t, size_t& b, size_t& e) {
      b = np * (size_t)t / (size_t)T;
      e = np * (size_t)(t + 1) / (size_t)T;
    };
    auto run = [&](auto&& fn) {
      std::vector<std::thread> pool;
      for (int t = 1; t < T; ++t) pool.emplace_back(fn, t);
      fn(0);
      for (auto& th : pool) th.join();
    };
    run([&](int t) {
      size_t b, e;
      chunk(t, b, e);
      Local& L = local[t];
      L.index.reserve((e - b) * 8 + 16);
      for (size_t i = b; i < e; ++i) {
        const Particle& q = particles[i];
        if (!q.enabled) continue;
        V3 gp = q.position * inv_h;
        int32_t bc[3];
        base_cell(gp, bc);
        V3 fx = gp - v3((double)bc[0], (double)bc[1], (double)bc[2]);
        double w[3][3];
        quadratic_weights(fx, w);
        M3 stress = kirchhoff_stress(q.material, q.deformation, q.affine, q.plastic);
        M3 momentum_affine = stress * (-dt * q.rest_volume * inv_d) + q.affine * q.mass;
        V3 mv = v3(q.mass * q.velocity.x, q.mass * q.velocity.y, q.mass * q.velocity.z);
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c)
            for (int z = 0; z < 3; ++z) {
              NodeKey k{bc[0] + a, bc[1] + c, bc[2] + z, keys[i]};
              double weight = w[0][a] * w[1][c] * w[2][z];
              V3 dpos = (v3((double)a, (double)c, (double)z) - fx) * h;
              auto it = L.index.find(k);
              uint32_t id;
              if (it == L.index.end()) {
                id = (uint32_t)L.nodes.size();
                L.index.emplace(k, id);
                L.keys.push_back(k);
                L.nodes.emplace_back();
              } else {
                id = it->second;
              }
              GridNode& n = L.nodes[id];
              n.mass += weight * q.mass;
              n.momentum = n.momentum + weight * (mv + momentum_affine * dpos);
            }
      }
    });
    // merge in thread order
    std::unordered_map<NodeKey, uint32_t, NodeKeyHash> index = std::move(local[0].index);
    std::vector<NodeKey> node_keys = std::move(local[0].keys);
    std::vector<GridNode> nodes = std::move(local[0].nodes);
    for (int t = 1; t < T; ++t) {
      Local& L = local[t];
      for (size_t j = 0; j < L.nodes.size(); ++j) {
        auto it = index.find(L.keys[j]);
        if (it == index.end()) {
          index.emplace(L.keys[j], (uint32_t)nodes.size());
          node_keys.push_back(L.keys[j]);
          nodes.push_back(L.nodes[j]);
        } else {
          GridNode& n = nodes[it->second];
          n.mass += L.nodes[j].mass;
          n.momentum = n.momentum + L.nodes[j].momentum;
        }
      }
      L = Local();
    }
    // grid update (node order is irrelevant without reactions)
    run([&](int t) {
      const size_t nn = nodes.size();
      for (size_t id = nn * (size_t)t / (size_t)T; id < nn * (size_t)(t + 1) / (size_t)T; ++id) {
        GridNode& n = nodes[id];
        if (n.mass <= 0.0) continue;
        n.velocity = n.momentum / n.mass + params.gravity * dt;
        const NodeKey& k = node_keys[id];
        V3 point = v3((double)k.x, (double)k.y, (double)k.z) * h;
        if (params.has_bounds)
          for (int a = 0; a < 3; ++a)
            if ((point[a] <= params.bmin[a] + h && n.velocity[a] < 0.0) ||
                (point[a] >= params.bmax[a] - h && n.velocity[a] > 0.0))
              n.velocity[a] = 0.0;
        for (const Obstacle& o : obstacles) {
          if (!may_contact(o, point, h)) continue;
          double dist;
          V3 normal;
          surface(o, point, dist, normal);
          if (dist <= h * (1.0 + 1e-5)) n.velocity = contact_velocity(o, point, n.velocity, normal);
        }
      }
    });
    std::vector<char> bad(T, 0);
    run([&](int t) {
      size_t b, e;
      chunk(t, b, e);
      for (size_t i = b; i < e; ++i) {
        Particle& q = particles[i];
        if (!q.enabled) {
          q.force = V3();
          continue;
        }
        if (q.fixed) {
          q.velocity = V3();
          q.affine = M3();
          q.force = V3();
          continue;
        }
        V3 gp = q.position * inv_h;
        int32_t bc[3];
        base_cell(gp, bc);
        V3 fx = gp - v3((double)bc[0], (double)bc[1], (double)bc[2]);
        double w[3][3];
        quadratic_weights(fx, w);
        V3 velocity;
        M3 affine;
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c)
            for (int z = 0; z < 3; ++z) {
              NodeKey k{bc[0] + a, bc[1] + c, bc[2] + z, keys[i]};
              double weight = w[0][a] * w[1][c] * w[2][z];
              V3 dpos = (v3((double)a, (double)c, (double)z) - fx) * h;
              auto it = index.find(k);
              V3 nv = it == index.end() ? V3() : nodes[it->second].velocity;
              velocity = velocity + nv * weight;
              affine = affine + outer(nv, dpos) * (weight * inv_d);
            }
        velocity = velocity + q.force / q.mass * dt;
        velocity = velocity * std::exp(-q.damping * dt);
        q.velocity = velocity;
        q.affine = affine;
        q.position = q.position + velocity * dt;
        q.deformation = (identity() + affine * dt) * q.deformation;
        project_deformation(q.material, q.deformation, q.plastic);
        q.force = V3();
        if (params.has_bounds)
          for (int a = 0; a < 3; ++a) {
            double lower = params.bmin[a] + q.radius, upper = params.bmax[a] - q.radius;
            if (q.position[a] < lower) {
              q.position[a] = lower;
              q.velocity[a] = std::fmax(q.velocity[a], 0.0);
            } else if (q.position[a] > upper) {
              q.position[a] = upper;
              q.velocity[a] = std::fmin(q.velocity[a], 0.0);
            }
          }
        for (const Obstacle& o : obstacles) {
          if (!may_contact(o, q.position, q.radius)) continue;
          double dist;
          V3 normal;
          surface(o, q.position, dist, normal);
          if (dist < q.radius) {
            q.position = q.position + normal * (q.radius - dist);
            q.velocity = contact_velocity(o, q.position, q.velocity, normal);
          }
        }
        if (!finite_particle_state(q)) bad[t] = 1;
      }
    });
    for (char c : bad)
      if (c) return NON_FINITE_STATE;
    substeps = substeps == UINT64_MAX ? substeps : substeps + 1;
    return OK;
  }
};

// ---------------------------------------------------------------------------
// FP32 restatement of the reference's resident GPU substep (gpu.wgsl:428-1019
// with the f32 packing of gpu.rs:1118-1297): the error floor that f32
// arithmetic alone leaves against the FP64 world above.  Every operation is
// an f32 operation in the WGSL's order, each rounded separately (the oracle
// is built with -ffp-contract=off).  Grid sums are added in particle order,
// one of the orders the reference's CAS loops may produce.  Scope: elastic
// (Newton polar, 10 iterations), neo-Hookean and fluid particles, gravity
// and walls; no obstacles or plastic materials (run_fp32 refuses them).
// ---------------------------------------------------------------------------
struct F3 {
  float x = 0, y = 0, z = 0;
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
inline F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 crossf(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// mat3x3<f32> as three columns
struct FM3 {
  F3 c[3];
};
inline F3 mulv(const FM3& m, F3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }
inline FM3 mulm(const FM3& a, const FM3& b) {
  FM3 r;
  for (int k = 0; k < 3; ++k) r.c[k] = mulv(a, b.c[k]);
  return r;
}
inline FM3 transposef(const FM3& m) {
  FM3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = m.c[j][i];
  return r;
}
inline FM3 addm(const FM3& a, const FM3& b) {
  FM3 r;
  for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] + b.c[k];
  return r;
}
inline FM3 subm(const FM3& a, const FM3& b) {
  FM3 r;
  for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] - b.c[k];
  return r;
}
inline FM3 scalem(const FM3& a, float s) {
  FM3 r;
  for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] * s;
  return r;
}
inline FM3 diagf(float d) {
  FM3 r;
  r.c[0].x = r.c[1].y = r.c[2].z = d;
  return r;
}
inline float detf(const FM3& m) { return dotf(m.c[0], crossf(m.c[1], m.c[2])); }

// gpu.wgsl:428-436
inline float qweight(float v, int o) {
  if (o == 0) return 0.5f * (1.5f - v) * (1.5f - v);
  if (o == 1) return 0.75f - (v - 1.0f) * (v - 1.0f);
  return 0.5f * (v - 0.5f) * (v - 0.5f);
}

// gpu.wgsl:451-465
inline bool regular_corotated_f(const FM3& F) {
  const F3 a = F.c[0], b = F.c[1], c = F.c[2];
  const float det = dotf(a, crossf(b, c));
  const float nrm = std::sqrt(dotf(a, a) + dotf(b, b) + dotf(c, c));
  if (std::fabs(det) < 1e-5f || nrm < 1e-5f || nrm > 20.0f) return false;
  const F3 i0 = crossf(b, c) / det, i1 = crossf(c, a) / det, i2 = crossf(a, b) / det;
  const float inrm = std::sqrt(dotf(i0, i0) + dotf(i1, i1) + dotf(i2, i2));
  return nrm * inrm <= 100.0f;
}

// gpu.wgsl:467-482
inline FM3 polar_rotation_f(const FM3& F) {
  FM3 r = F;
  for (int it = 0; it < 10; ++it) {
    const F3 a = r.c[0], b = r.c[1], c = r.c[2];
    const float d = dotf(a, crossf(b, c));
    FM3 inv_t;
    inv_t.c[0] = crossf(b, c) / d;
    inv_t.c[1] = crossf(c, a) / d;
    inv_t.c[2] = crossf(a, b) / d;
    r = scalem(addm(r, inv_t), 0.5f);
  }
  return r;
}

struct P32 {
  F3 x, v, force;
  float m = 0, vol = 0, radius = 0, damping = 0;
  FM3 F, C;
  int code = 0;  // gpu.rs:1171-1276: 0 host stress, 1 neo-Hookean, 2 fluid, 3 corotated
  float lambda = 0, mu = 0, bulk = 0, gamma = 0, viscosity = 0, tensile = 0;
  FM3 host_stress;  // code 0: the f64 stress times -dt V0 inv_d, rounded to f32
  int32_t base[3] = {0, 0, 0};
  uint16_t color = 0;
  bool enabled = true, fixed = false;
};

inline float to_f32(double v) { return (float)v; }
inline FM3 to_fm3(const M3& m) {
  FM3 r;
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k) r.c[c][k] = (float)m.a[k][c];
  return r;
}
inline M3 to_m3(const FM3& f) {
  M3 r;
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k) r.a[k][c] = (double)f.c[c][k];
  return r;
}

// gpu.wgsl:638-701
FM3 particle_stress_f(const P32& q, float dt, float inv_d) {
  const FM3& F = q.F;
  const float j = detf(F);
  FM3 stress;
  if (q.code == 0) return q.host_stress;
  if (q.code == 1) {
    const float safe_j = std::fmax(j, 1e-10f);
    const float iso = q.lambda * std::log(safe_j) - q.mu;
    stress = addm(scalem(mulm(F, transposef(F)), q.mu), diagf(iso));
  } else if (q.code == 2) {
    const float safe_j = std::fmax(j, 1e-6f);
    float pressure = -q.bulk * q.tensile * (safe_j - 1.0f);
    if (safe_j <= 1.0f) pressure = q.bulk * (std::pow(safe_j, -q.gamma) - 1.0f);
    const FM3 sr = scalem(addm(q.C, transposef(q.C)), 0.5f);
    const float trace_third = (sr.c[0].x + sr.c[1].y + sr.c[2].z) / 3.0f;
    const float diagonal = -2.0f * q.viscosity * safe_j * trace_third - pressure * safe_j;
    stress = addm(scalem(sr, 2.0f * q.viscosity * safe_j), diagf(diagonal));
  } else {
    const FM3 R = polar_rotation_f(F);
    const float iso = q.lambda * 1.0f * (j - 1.0f) * j;
    stress = addm(scalem(mulm(subm(F, R), transposef(F)), 2.0f * q.mu * 1.0f), diagf(iso));
  }
  return scalem(stress, -dt * q.vol * inv_d);
}

struct NodeKey32 {
  int32_t x, y, z;
  uint32_t color;
  bool operator==(const NodeKey32& o) const { return x == o.x && y == o.y && z == o.z && color == o.color; }
};
struct NodeKey32Hash {
  size_t operator()(const NodeKey32& k) const {
    uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.color * 0x27D4EB2F165667C5ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};
struct Node32 {
  float m = 0;
  F3 p, v;
};

// Packs the world (gpu.rs:1118-1297), runs `steps` resident substeps of dt in
// f32 and writes the state back as f64 (insertion order; disabled particles
// keep theirs).  Returns INVALID_INPUT for what the restatement leaves out.
int run_fp32(const World& w, double dt, uint32_t steps, double* pos, double* vel, double* Fo, double* Co) {
  if (!w.obstacles.empty()) return INVALID_INPUT;
  const Params& pr = w.params;
  const float h = to_f32(pr.cell_width), step = to_f32(dt), inv_d = to_f32(4.0 / (pr.cell_width * pr.cell_width));
  const float inv_h = 1.0f / h;
  const F3 g = f3(to_f32(pr.gravity.x), to_f32(pr.gravity.y), to_f32(pr.gravity.z));
  const F3 bmin = f3(to_f32(pr.bmin.x), to_f32(pr.bmin.y), to_f32(pr.bmin.z));
  const F3 bmax = f3(to_f32(pr.bmax.x), to_f32(pr.bmax.y), to_f32(pr.bmax.z));
  const size_t n = w.particles.size();
  std::vector<P32> ps(n);
  for (size_t i = 0; i < n; ++i) {
    const Particle& q = w.particles[i];
    P32& r = ps[i];
    r.x = f3(to_f32(q.position.x), to_f32(q.position.y), to_f32(q.position.z));
    r.v = f3(to_f32(q.velocity.x), to_f32(q.velocity.y), to_f32(q.velocity.z));
    r.force = f3(to_f32(q.force.x), to_f32(q.force.y), to_f32(q.force.z));
    r.m = q.enabled ? to_f32(q.mass) : 0.0f;
    r.vol = to_f32(q.rest_volume);
    r.radius = to_f32(q.radius);
    r.damping = to_f32(q.damping);
    r.F = to_fm3(q.deformation);
    r.C = to_fm3(q.affine);
    r.color = q.transfer_color;
    r.enabled = q.enabled;
    r.fixed = q.fixed;
    for (int a = 0; a < 3; ++a) r.base[a] = (int32_t)std::floor(r.x[a] * inv_h - 0.5f);
    const Material& mt = q.material;
    double lam = 0, mu = 0;
    switch (mt.kind) {
      case 0:  // LinearElastic (corotated)
        lame(mt.p[0], mt.p[1], lam, mu);
        r.code = 3;
        r.lambda = to_f32(lam);
        r.mu = to_f32(mu);
        break;
      case 1:  // NeoHookean
        lame(mt.p[0], mt.p[1], lam, mu);
        r.code = 1;
        r.lambda = to_f32(lam);
        r.mu = to_f32(mu);
        break;
      case 3:  // Fluid
        r.code = 2;
        r.bulk = to_f32(mt.p[0]);
        r.gamma = to_f32(mt.p[1]);
        r.viscosity = to_f32(mt.p[2]);
        r.tensile = to_f32(mt.p[3]);
        break;
      default:
        if (q.enabled) return INVALID_INPUT;
    }
  }
  std::unordered_map<NodeKey32, uint32_t, NodeKey32Hash> index;
  std::vector<NodeKey32> keys;
  std::vector<Node32> nodes;
  for (uint32_t sstep = 0; sstep < steps; ++sstep) {
    index.clear();
    keys.clear();
    nodes.clear();
    index.reserve(n * 8 + 16);
    // p2g (gpu.wgsl:711-770)
    for (size_t i = 0; i < n; ++i) {
      P32& q = ps[i];
      if (q.m <= 0.0f) continue;
      if (q.code == 3 && !regular_corotated_f(q.F)) {
        // the reference's direct path computes this stress on the host in f64
        const Particle& orig = w.particles[i];
        const M3 s64 = corotated_stress(to_m3(q.F), orig.material.p[0], orig.material.p[1], 1.0) *
                       (-dt * orig.rest_volume * (double)inv_d);
        q.host_stress = to_fm3(s64);
        q.code = 0;
      }
      const F3 gp = q.x * inv_h;
      const F3 fx = gp - f3((float)q.base[0], (float)q.base[1], (float)q.base[2]);
      const FM3 S = particle_stress_f(q, step, inv_d);
      if (q.code == 0) q.code = 3;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            const float weight = qweight(fx.x, a) * qweight(fx.y, b) * qweight(fx.z, c);
            const F3 dpos = (f3((float)a, (float)b, (float)c) - fx) * h;
            const F3 am = (q.C.c[0] * q.m + S.c[0]) * dpos.x + (q.C.c[1] * q.m + S.c[1]) * dpos.y +
                          (q.C.c[2] * q.m + S.c[2]) * dpos.z;
            const F3 mom = weight * (q.m * q.v + am);
            const NodeKey32 k{q.base[0] + a, q.base[1] + b, q.base[2] + c, q.color};
            auto it = index.find(k);
            uint32_t id;
            if (it == index.end()) {
              id = (uint32_t)nodes.size();
              index.emplace(k, id);
              keys.push_back(k);
              nodes.emplace_back();
            } else {
              id = it->second;
            }
            Node32& nd = nodes[id];
            nd.m = nd.m + weight * q.m;
            nd.p = nd.p + mom;
          }
    }
    // update_grid (gpu.wgsl:772-819)
    for (size_t id = 0; id < nodes.size(); ++id) {
      Node32& nd = nodes[id];
      if (nd.m <= 0.0f) continue;
      F3 v = nd.p / nd.m + g * step;
      const F3 position = f3((float)keys[id].x, (float)keys[id].y, (float)keys[id].z) * h;
      if (pr.has_bounds)
        for (int a = 0; a < 3; ++a)
          if ((position[a] <= bmin[a] + h && v[a] < 0.0f) || (position[a] >= bmax[a] - h && v[a] > 0.0f)) v[a] = 0.0f;
      nd.v = v;
    }
    // g2p (gpu.wgsl:821-1019)
    for (size_t i = 0; i < n; ++i) {
      P32& q = ps[i];
      if (q.m <= 0.0f) continue;
      if (q.fixed) {
        q.v = F3();
        q.force = F3();
        q.C = FM3();
        continue;
      }
      const F3 gp = q.x * inv_h;
      const F3 fx = gp - f3((float)q.base[0], (float)q.base[1], (float)q.base[2]);
      F3 velocity, a0, a1, a2;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            const float weight = qweight(fx.x, a) * qweight(fx.y, b) * qweight(fx.z, c);
            const F3 dpos = (f3((float)a, (float)b, (float)c) - fx) * h;
            auto it = index.find(NodeKey32{q.base[0] + a, q.base[1] + b, q.base[2] + c, q.color});
            const F3 nv = it == index.end() ? F3() : nodes[it->second].v;
            velocity = velocity + nv * weight;
            a0 = a0 + nv * (dpos.x * weight * inv_d);
            a1 = a1 + nv * (dpos.y * weight * inv_d);
            a2 = a2 + nv * (dpos.z * weight * inv_d);
          }
      velocity = velocity + q.force / q.m * step;
      velocity = velocity * std::exp(-q.damping * step);
      F3 position = q.x + velocity * step;
      if (pr.has_bounds)
        for (int a = 0; a < 3; ++a) {
          const float lower = bmin[a] + q.radius, upper = bmax[a] - q.radius;
          if (position[a] < lower) {
            position[a] = lower;
            velocity[a] = std::fmax(velocity[a], 0.0f);
          } else if (position[a] > upper) {
            position[a] = upper;
            velocity[a] = std::fmin(velocity[a], 0.0f);
          }
        }
      FM3 affine;
      affine.c[0] = a0;
      affine.c[1] = a1;
      affine.c[2] = a2;
      FM3 def = mulm(addm(diagf(1.0f), scalem(affine, step)), q.F);
      if (q.code == 2) def = diagf(std::pow(std::fmax(detf(def), 1e-9f), 1.0f / 3.0f));
      q.x = position;
      q.v = velocity;
      q.force = F3();
      q.C = affine;
      q.F = def;
      for (int a = 0; a < 3; ++a) {
        const float nc = std::floor(position[a] * inv_h - 0.5f);
        if (std::isfinite(nc) && nc >= -2147483648.0f && nc < 2147483648.0f) q.base[a] = (int32_t)nc;
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const Particle& orig = w.particles[i];
    const P32& q = ps[i];
    const bool moved = orig.enabled;
    for (int a = 0; a < 3; ++a) {
      pos[3 * i + a] = moved ? (double)q.x[a] : orig.position[a];
      vel[3 * i + a] = moved ? (double)q.v[a] : orig.velocity[a];
    }
    const M3 F = moved ? to_m3(q.F) : orig.deformation, C = moved ? to_m3(q.C) : orig.affine;
    store_cm(F, Fo + 9 * i);
    store_cm(C, Co + 9 * i);
  }
  return OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI for the ctypes test harness (oracle/oracle.py).
// ---------------------------------------------------------------------------
extern "C" {

struct orc_params {
  double cell_width;
  double gravity[3];
  double max_substep;
  int32_t has_bounds;
  int32_t use_cpic;
  double bound_min[3];
  double bound_max[3];
};

// Same POD layout as tmpm_obstacle in include/tessera_mpm.h.
struct orc_obstacle {
  int32_t shape;
  int32_t boundary;
  int32_t cpic_group;
  uint32_t plane_count;
  double center[3];
  double orientation[4];
  double linear_velocity[3];
  double angular_velocity[3];
  double friction;
  double dims[4];
  double vertices[9];
  const double* planes;
};

static Obstacle to_obstacle(const orc_obstacle& s) {
  Obstacle o;
  o.shape = s.shape;
  o.boundary = s.boundary;
  o.cpic_group = s.cpic_group;
  o.center = v3(s.center[0], s.center[1], s.center[2]);
  o.orientation.i = s.orientation[0];
  o.orientation.j = s.orientation[1];
  o.orientation.k = s.orientation[2];
  o.orientation.w = s.orientation[3];
  o.linear_velocity = v3(s.linear_velocity[0], s.linear_velocity[1], s.linear_velocity[2]);
  o.angular_velocity = v3(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2]);
  o.friction = s.friction;
  for (int i = 0; i < 4; ++i) o.dims[i] = s.dims[i];
  for (int v = 0; v < 3; ++v) o.vertices[v] = v3(s.vertices[3 * v], s.vertices[3 * v + 1], s.vertices[3 * v + 2]);
  if (s.shape == CONVEX && s.planes) {
    o.planes.assign(s.planes, s.planes + 4 * (size_t)s.plane_count);
    o.bound_radius = s.dims[0];
  }
  return o;
}

void* orc_world_new(const orc_params* p) {
  World* w = new World();
  w->params.cell_width = p->cell_width;
  w->params.gravity = v3(p->gravity[0], p->gravity[1], p->gravity[2]);
  w->params.max_substep = p->max_substep;
  w->params.has_bounds = p->has_bounds != 0;
  w->params.bmin = v3(p->bound_min[0], p->bound_min[1], p->bound_min[2]);
  w->params.bmax = v3(p->bound_max[0], p->bound_max[1], p->bound_max[2]);
  w->use_cpic = p->use_cpic != 0;
  if (!params_valid(w->params)) {
    delete w;
    return nullptr;
  }
  return w;
}

void orc_world_free(void* h) { delete static_cast<World*>(h); }

// Appends a validated batch (world.rs:225-265).  is_initial != 0 assigns chunk
// 0 (MpmWorld::new); otherwise the next chunk id is consumed.  Returns status.
int orc_world_add(void* h, uint64_t n, int32_t is_initial, const double* pos, const double* vel, const double* F,
                  const double* C, const double* force, const double* mass, const double* rest_volume,
                  const double* radius, const double* damping, const int32_t* mat_kind, const double* mat_params,
                  const double* plastic, const uint16_t* color, const uint8_t* enabled, const uint8_t* fixed,
                  uint64_t* chunk_out) {
  World* w = static_cast<World*>(h);
  if (!is_initial && n == 0) return INVALID_INPUT;
  std::vector<Particle> batch(n);
  for (uint64_t i = 0; i < n; ++i) {
    Particle& q = batch[i];
    q.position = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    q.velocity = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    q.deformation = load_cm(F + 9 * i);
    q.affine = load_cm(C + 9 * i);
    q.force = v3(force[3 * i], force[3 * i + 1], force[3 * i + 2]);
    q.mass = mass[i];
    q.rest_volume = rest_volume[i];
    q.radius = radius[i];
    q.damping = damping[i];
    q.material.kind = mat_kind[i];
    for (int k = 0; k < 5; ++k) q.material.p[k] = mat_params[5 * i + k];
    q.plastic.plastic_det = plastic[3 * i];
    q.plastic.hardening = plastic[3 * i + 1];
    q.plastic.log_volume_gain = plastic[3 * i + 2];
    q.transfer_color = color[i];
    q.enabled = enabled[i] != 0;
    q.fixed = fixed[i] != 0;
    if (!valid_particle(q)) return INVALID_INPUT;
  }
  uint64_t chunk = 0;
  if (!is_initial) {
    if (w->next_chunk_id == UINT64_MAX) return PARTICLE_CAPACITY;
    chunk = w->next_chunk_id++;
  }
  for (Particle& q : batch) q.chunk = chunk;
  w->particles.insert(w->particles.end(), batch.begin(), batch.end());
  if (chunk_out) *chunk_out = chunk;
  return OK;
}

// world.rs:268-277
int orc_world_remove_chunk(void* h, uint64_t chunk, uint64_t* removed) {
  World* w = static_cast<World*>(h);
  size_t before = w->particles.size();
  w->particles.erase(std::remove_if(w->particles.begin(), w->particles.end(),
                                    [&](const Particle& q) { return q.chunk == chunk; }),
                     w->particles.end());
  *removed = before - w->particles.size();
  return *removed == 0 ? UNKNOWN_CHUNK : OK;
}

// world.rs:280-286
int orc_world_set_obstacles(void* h, uint64_t n, const orc_obstacle* obs) {
  World* w = static_cast<World*>(h);
  std::vector<Obstacle> list;
  for (uint64_t i = 0; i < n; ++i) {
    list.push_back(to_obstacle(obs[i]));
    if (!obstacle_valid(list.back())) return INVALID_INPUT;
  }
  w->obstacles = list;
  return OK;
}

void orc_world_set_use_cpic(void* h, int32_t on) { static_cast<World*>(h)->use_cpic = on != 0; }

int orc_world_step(void* h, double dt) { return static_cast<World*>(h)->advance(dt, nullptr); }

// world.rs:314-330; out receives 6 doubles per obstacle (linear, angular).
int orc_world_step_with_reactions(void* h, double dt, double* out) {
  World* w = static_cast<World*>(h);
  std::vector<Reaction> r(w->obstacles.size());
  int st = w->advance(dt, &r);
  if (st != OK) return st;
  for (size_t i = 0; i < r.size(); ++i) {
    V3 l = r[i].linear, a = r[i].angular;
    if (!finite3(l) || !finite3(a)) return NON_FINITE_STATE;
    out[6 * i] = l.x; out[6 * i + 1] = l.y; out[6 * i + 2] = l.z;
    out[6 * i + 3] = a.x; out[6 * i + 4] = a.y; out[6 * i + 5] = a.z;
  }
  return OK;
}

// Fixed-size substeps (the resident session's semantics: each substep uses
// exactly dt, no CFL subdivision).
int orc_world_substeps_fixed(void* h, double dt, uint32_t steps) {
  World* w = static_cast<World*>(h);
  for (uint32_t s = 0; s < steps; ++s) {
    int st = w->substep(dt, nullptr);
    if (st != OK) return st;
  }
  return OK;
}

// Fixed-size substeps over `threads` host threads (CPU baseline timing).
int orc_world_substeps_fixed_threads(void* h, double dt, uint32_t steps, int32_t threads) {
  World* w = static_cast<World*>(h);
  if (!std::isfinite(dt) || dt <= 0.0) return INVALID_INPUT;
  for (uint32_t k = 0; k < steps; ++k) {
    int st = w->substep_threads(dt, threads);
    if (st != OK) return st;
  }
  return OK;
}

double orc_world_stable_timestep(void* h) { return static_cast<World*>(h)->stable_timestep(); }

// The FP32 restatement of `steps` resident substeps of dt from the world's
// current state (the world itself is not changed): positions, velocities, F
// and C in insertion order.
int orc_world_fp32_substeps(void* h, double dt, uint32_t steps, double* pos, double* vel, double* F, double* C) {
  return run_fp32(*static_cast<World*>(h), dt, steps, pos, vel, F, C);
}
uint64_t orc_world_count(void* h) { return static_cast<World*>(h)->particles.size(); }
uint64_t orc_world_substeps(void* h) { return static_cast<World*>(h)->substeps; }

void orc_world_read(void* h, double* pos, double* vel, double* F, double* C, double* force, double* plastic,
                    uint64_t* chunk) {
  World* w = static_cast<World*>(h);
  for (size_t i = 0; i < w->particles.size(); ++i) {
    const Particle& q = w->particles[i];
    pos[3 * i] = q.position.x; pos[3 * i + 1] = q.position.y; pos[3 * i + 2] = q.position.z;
    vel[3 * i] = q.velocity.x; vel[3 * i + 1] = q.velocity.y; vel[3 * i + 2] = q.velocity.z;
    store_cm(q.deformation, F + 9 * i);
    store_cm(q.affine, C + 9 * i);
    force[3 * i] = q.force.x; force[3 * i + 1] = q.force.y; force[3 * i + 2] = q.force.z;
    plastic[3 * i] = q.plastic.plastic_det;
    plastic[3 * i + 1] = q.plastic.hardening;
    plastic[3 * i + 2] = q.plastic.log_volume_gain;
    if (chunk) chunk[i] = q.chunk;
  }
}

// Overwrite state arrays in place (same layout as orc_world_read); used by
// tests that perturb one field.
void orc_world_write(void* h, const double* pos, const double* vel, const double* F, const double* C,
                     const double* force) {
  World* w = static_cast<World*>(h);
  for (size_t i = 0; i < w->particles.size(); ++i) {
    Particle& q = w->particles[i];
    q.position = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    q.velocity = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    q.deformation = load_cm(F + 9 * i);
    q.affine = load_cm(C + 9 * i);
    q.force = v3(force[3 * i], force[3 * i + 1], force[3 * i + 2]);
  }
}

// One P2G pass only: returns the number of grid nodes and writes the total
// node mass and momentum (for partition-of-unity / conservation checks).
uint64_t orc_world_p2g_totals(void* h, double dt, double* mass_total, double* momentum_total) {
  World* w = static_cast<World*>(h);
  World copy = *w;
  const double hh = copy.params.cell_width, inv_h = 1.0 / hh, inv_d = 4.0 * inv_h * inv_h;
  std::unordered_map<NodeKey, GridNode, NodeKeyHash> grid;
  for (const Particle& q : copy.particles) {
    if (!q.enabled) continue;
    V3 gp = q.position * inv_h;
    int32_t b[3];
    base_cell(gp, b);
    V3 fx = gp - v3((double)b[0], (double)b[1], (double)b[2]);
    double wt[3][3];
    quadratic_weights(fx, wt);
    M3 stress = kirchhoff_stress(q.material, q.deformation, q.affine, q.plastic);
    M3 ma = stress * (-dt * q.rest_volume * inv_d) + q.affine * q.mass;
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) {
          NodeKey k{b[0] + a, b[1] + c, b[2] + e, copy.transfer_key(q)};
          double weight = wt[0][a] * wt[1][c] * wt[2][e];
          V3 dpos = (v3(a, c, e) - fx) * hh;
          GridNode& n = grid[k];
          n.mass += weight * q.mass;
          n.momentum = n.momentum + weight * (q.velocity * q.mass + ma * dpos);
        }
  }
  double m = 0;
  V3 p;
  for (auto& kv : grid) {
    m += kv.second.mass;
    p = p + kv.second.momentum;
  }
  *mass_total = m;
  momentum_total[0] = p.x; momentum_total[1] = p.y; momentum_total[2] = p.z;
  return grid.size();
}

// --- materials -------------------------------------------------------------
void orc_kirchhoff_stress(int32_t kind, const double* mp, const double* F, const double* C, const double* plastic,
                          double* out) {
  Material m;
  m.kind = kind;
  for (int k = 0; k < 5; ++k) m.p[k] = mp[k];
  Plastic pl;
  pl.plastic_det = plastic[0]; pl.hardening = plastic[1]; pl.log_volume_gain = plastic[2];
  store_cm(kirchhoff_stress(m, load_cm(F), load_cm(C), pl), out);
}
void orc_project_deformation(int32_t kind, const double* mp, const double* F, const double* plastic, double* F_out,
                             double* plastic_out) {
  Material m;
  m.kind = kind;
  for (int k = 0; k < 5; ++k) m.p[k] = mp[k];
  Plastic pl;
  pl.plastic_det = plastic[0]; pl.hardening = plastic[1]; pl.log_volume_gain = plastic[2];
  M3 f = load_cm(F);
  project_deformation(m, f, pl);
  store_cm(f, F_out);
  plastic_out[0] = pl.plastic_det; plastic_out[1] = pl.hardening; plastic_out[2] = pl.log_volume_gain;
}
double orc_timestep_bound(int32_t kind, const double* mp, double density, const double* v, double h, double det) {
  Material m;
  m.kind = kind;
  for (int k = 0; k < 5; ++k) m.p[k] = mp[k];
  return timestep_bound(m, density, v3(v[0], v[1], v[2]), h, det);
}
int32_t orc_material_valid(int32_t kind, const double* mp) {
  Material m;
  m.kind = kind;
  for (int k = 0; k < 5; ++k) m.p[k] = mp[k];
  return material_valid(m) ? 1 : 0;
}
void orc_svd(const double* F, double* U, double* s, double* V) {
  M3 u, v;
  V3 sv;
  svd3(load_cm(F), u, sv, v);
  store_cm(u, U);
  store_cm(v, V);
  s[0] = sv.x; s[1] = sv.y; s[2] = sv.z;
}
double orc_determinant(const double* F) { return determinant(load_cm(F)); }

// --- obstacles -------------------------------------------------------------
void orc_obstacle_surface(const orc_obstacle* o, const double* p, double* dist, double* normal) {
  Obstacle ob = to_obstacle(*o);
  V3 n;
  surface(ob, v3(p[0], p[1], p[2]), *dist, n);
  normal[0] = n.x; normal[1] = n.y; normal[2] = n.z;
}
void orc_obstacle_contact_velocity(const orc_obstacle* o, const double* p, const double* v, const double* n,
                                   double* out) {
  Obstacle ob = to_obstacle(*o);
  V3 r = contact_velocity(ob, v3(p[0], p[1], p[2]), v3(v[0], v[1], v[2]), v3(n[0], n[1], n[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int32_t orc_obstacle_may_contact(const orc_obstacle* o, const double* p, double margin) {
  return may_contact(to_obstacle(*o), v3(p[0], p[1], p[2]), margin) ? 1 : 0;
}
int32_t orc_obstacle_valid(const orc_obstacle* o) { return obstacle_valid(to_obstacle(*o)) ? 1 : 0; }

// --- binning (f32 cell-key rule) ----------------------------------------------
// The HIP backend keys particles on f32 positions exactly like the reference
// GPU host packing (gpu.rs:1118-1154): h32 = (f32)h, inv_h = 1/h32 in f32,
// cell = floor(x32 * inv_h - 0.5) with each operation rounded to f32 and no
// fused multiply-add.  Returns 1 and writes the cell when finite and within
// [i32::MIN, i32::MAX - 2] on every axis; 0 otherwise.
int32_t orc_cell_f32(const float* x, float h32, int32_t* cell) {
  volatile float inv_h = 1.0f / h32;
  for (int a = 0; a < 3; ++a) {
    volatile float gp = x[a] * inv_h;
    volatile float shifted = gp - 0.5f;
    float c = std::floor((float)shifted);
    if (!std::isfinite(c) || (double)c < -2147483648.0 || (double)c > 2147483645.0) return 0;
    cell[a] = (int32_t)c;
  }
  return 1;
}

// Live grid-node set: the union over active particles of their 27-node
// stencils keyed (cell, region) (gpu_topology.wgsl:18-66; world.rs:378-399).
// `region` carries the transfer colour (and, in batches, the scene).  Writes
// up to `cap` sorted 4-int keys (x, y, z, region) and returns the set size, or
// UINT64_MAX if some active particle has an unrepresentable cell.
//
// The union of the stencils {c + (a, b, e) : a, b, e in 0..2} is the distinct
// base cells dilated by the box [0, 2]^3, which is three one-axis dilations
// by [0, 2] in turn (a box is the product of its edges).  A call with out ==
// NULL (the size query) keeps its result for this thread, and the next call
// with the same inputs copies it instead of building the set again.
uint64_t orc_node_set(uint64_t n, const float* pos, const uint8_t* active, const uint32_t* region, float h32,
                      int32_t* out, uint64_t cap) {
  using Key = std::array<int32_t, 4>;
  struct Memo {
    uint64_t n = 0;
    const float* pos = nullptr;
    const uint8_t* active = nullptr;
    const uint32_t* region = nullptr;
    float h = 0.0f;
    std::vector<Key> keys;
  };
  thread_local Memo memo;
  const bool same = memo.pos != nullptr && memo.n == n && memo.pos == pos && memo.active == active &&
                    memo.region == region && memo.h == h32;
  if (!(out && same)) {
    memo = Memo();
    std::vector<Key> cells;
    cells.reserve(n);
    for (uint64_t i = 0; i < n; ++i) {
      if (!active[i]) continue;
      int32_t c[3];
      if (!orc_cell_f32(pos + 3 * i, h32, c)) return UINT64_MAX;
      cells.push_back({c[0], c[1], c[2], (int32_t)region[i]});
    }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    for (int axis = 2; axis >= 0; --axis) {
      std::vector<Key> grown;
      grown.reserve(cells.size() * 3);
      for (const Key& c : cells)
        for (int d = 0; d < 3; ++d) {
          Key k = c;
          k[axis] += d;
          grown.push_back(k);
        }
      std::sort(grown.begin(), grown.end());
      grown.erase(std::unique(grown.begin(), grown.end()), grown.end());
      cells.swap(grown);
    }
    memo.keys.swap(cells);
    memo.n = n;
    memo.pos = pos;
    memo.active = active;
    memo.region = region;
    memo.h = h32;
  }
  const std::vector<Key>& keys = memo.keys;
  for (uint64_t i = 0; out && i < keys.size() && i < cap; ++i)
    for (int k = 0; k < 4; ++k) out[4 * i + k] = keys[i][k];
  const uint64_t size = keys.size();
  if (out) memo = Memo();
  return size;
}

// Particle-block partition: block = cell >> 2 (floor division by 4) per axis,
// keyed with the region.  Writes the sorted unique block keys (up to cap) and
// for every particle the index of its block in that sorted list (-1 if
// inactive).  Returns the number of blocks (UINT64_MAX on a bad cell).
uint64_t orc_block_partition(uint64_t n, const float* pos, const uint8_t* active, const uint32_t* region, float h32,
                             int32_t* blocks_out, uint64_t cap, int64_t* particle_block) {
  std::vector<std::array<int32_t, 4>> keys(n);
  std::vector<std::array<int32_t, 4>> uniq;
  for (uint64_t i = 0; i < n; ++i) {
    if (!active[i]) continue;
    int32_t c[3];
    if (!orc_cell_f32(pos + 3 * i, h32, c)) return UINT64_MAX;
    keys[i] = {c[0] >> 2, c[1] >> 2, c[2] >> 2, (int32_t)region[i]};
    uniq.push_back(keys[i]);
  }
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  for (uint64_t i = 0; i < uniq.size() && i < cap; ++i)
    for (int k = 0; k < 4; ++k) blocks_out[4 * i + k] = uniq[i][k];
  for (uint64_t i = 0; i < n; ++i) {
    if (!active[i]) {
      particle_block[i] = -1;
      continue;
    }
    particle_block[i] = std::lower_bound(uniq.begin(), uniq.end(), keys[i]) - uniq.begin();
  }
  return uniq.size();
}

}  // extern "C"
