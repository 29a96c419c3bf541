"""finitevolumegpu_amd: MI355X (gfx950) flux-and-update engine for the
dimensionally split finite-volume schemes of FiniteVolumeGPU.

Public modules mirror the reference package ``GPUSimulators``:
``HipContext`` (CudaContext), ``Common``, ``Simulator``, ``KP07_dimsplit``,
``HLL2``, ``WAF``, ``EE2D_KP07_dimsplit``, ``MPISimulator``, ``Autotuner``;
plus ``Ensemble`` (batched many-grid engine) and ``Serving`` (continuous
batching).  The compute path is libfvmi.so (include/fvmi.h); importing a
simulator without the built library raises.
"""

__all__ = ["HipContext", "Common", "Simulator", "KP07_dimsplit", "HLL2", "WAF", "EE2D_KP07_dimsplit"]
