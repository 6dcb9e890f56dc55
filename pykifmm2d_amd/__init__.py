"""pykifmm2d_amd -- 2-D kernel-independent FMM for one AMD Instinct MI355X (gfx950).

Drop-in modules for the reference pykifmm2d entry points:
  pykifmm2d_amd.fmm        on_the_fly_fmm, fmm_planner, planned_fmm, Get_Kernel_Functions, ...
  pykifmm2d_amd.class_fmm  FMM (build_tree / precompute / build_expansions / evaluate_to_points)
  pykifmm2d_amd.svd_fmm    the same entry points under the reference's svd_fmm name
  pykifmm2d_amd.tree       Tree
  pykifmm2d_amd.utils      random, random2 (the reference's misc/utils.py)
Kernels are descriptors (pykifmm2d_amd.kernels) or the reference's own radial Kernel_Eval
functions, which are tabulated for the GPU.
"""
from . import kernels  # noqa: F401
from . import utils  # noqa: F401

__all__ = ["kernels", "utils", "Tree", "on_the_fly_fmm"]


def __getattr__(name):
    if name == "Tree":
        from .tree import Tree
        return Tree
    if name == "on_the_fly_fmm":
        from .fmm import on_the_fly_fmm
        return on_the_fly_fmm
    raise AttributeError(name)
