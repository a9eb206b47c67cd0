"""TEST INFRASTRUCTURE ONLY: the FP64 CPU oracle of the MLS-MPM substep.

Only tests/, the smoke check and the benchmark's CPU-baseline leg may import
this package; the product (tessera_amd) never does.
"""
