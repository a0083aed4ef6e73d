"""cumulus3d_amd — MI355X (gfx950) PatchStereo depth-map engine.

A drop-in for the PatchStereo hot path of cumulus3d behind its PatchMatchBackend boundary
(crates/dense/src/kernel.rs:191-245): hand-written HIP kernels for gfx950 under a C ABI
(include/c3d_pm.h), bound here with the reference's interface.

Modules:
``abi`` (ctypes mirror of the C ABI and parameter records), ``pm`` (backend and session, with the
device-resident compute_depth_maps schedule as ``Session.densify``), ``init_net`` (the learned
initialiser), ``fusion`` (agreement fusion and cloud comparison), ``neighbors`` (source-view
selection and depth ranges), ``autotune`` (the learned-schedule autotuner), ``train_init``
(training of the initialiser), ``scenes`` (synthetic scenes and kernel inputs), ``render``
(the terrain renderer on the device), ``profsum`` (rocprofv3 summaries), ``build`` (in-tree
compilation).
"""
from . import abi  # noqa: F401

__all__ = ["abi", "pm", "init_net", "fusion", "neighbors", "autotune", "train_init", "scenes", "render", "profsum", "build"]
