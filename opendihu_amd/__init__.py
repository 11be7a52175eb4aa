"""MI355X-native FastMonodomainSolver fibre path (OpenDiHu's monodomain hot path).

* ``opendihu_amd.fibers.FiberBatch`` - handle over fibres on the GPU (C ABI in
  include/opendihu_fms.h, kernels in opendihu_amd/csrc).
* ``opendihu_amd.fast_monodomain_solver.FastMonodomainSolver`` - the reference's
  settings-dict front end (RepeatedCall / MultipleInstances / StrangSplitting /
  Heun / CellML / ImplicitEuler|CrankNicolson keys).
* ``opendihu_amd.cellml`` - OpenCOR C model parser and HIP code generator.
"""
__all__ = ["fibers", "build", "cellml"]
