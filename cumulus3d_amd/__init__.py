"""cumulus3d_amd — MI355X (gfx950) PatchStereo depth-map engine.

A drop-in for the PatchStereo hot path of cumulus3d behind its PatchMatchBackend boundary
(crates/dense/src/kernel.rs:191-245): hand-written HIP kernels for gfx950 under a C ABI
(include/c3d_pm.h), bound here with the reference's interface.

Modules: ``abi`` (ctypes mirror of the C ABI and parameter records), ``pm`` (backend/session),
``scenes`` (synthetic scenes and kernel inputs), ``densify`` (the reference's compute_depth_maps
driver over a session), ``build`` (in-tree compilation).
"""
from . import abi  # noqa: F401

__all__ = ["abi", "pm", "scenes", "densify", "build"]
