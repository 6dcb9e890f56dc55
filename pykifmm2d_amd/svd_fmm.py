"""Drop-in for the reference's svd_fmm module, whose on_the_fly_fmm the reference's examples call
(examples/laplace.py:106, examples/modified_helmholtz.py:116).

svd_fmm.py differs from fmm.py in how it applies the inverse of a level's equivalent-to-check
matrix: the matrix is circulant (both circles carry p equispaced points at the same angles), so
svd_fmm divides by its eigenvalues in Fourier space (svd_fmm.py:239-242, 282-289) where fmm.py
LU-solves it (fmm.py:279, 363).  Both are the exact inverse of the same matrix.  libkifmm2d
applies it in the factored form F diag(1/lambda) F^T for every entry point, so the entry points
here are fmm's: on_the_fly_fmm (svd_fmm.py:191-237), fmm_planner (svd_fmm.py:583-766),
planned_fmm (svd_fmm.py:794-875), with the same factories.
"""
from .fmm import (FMM_Plan, Get_Kernel_Functions, Kernel_Apply, Kernel_Form, fmm_planner,  # noqa: F401
                  get_level_information, get_print_function, on_the_fly_fmm, planned_fmm,
                  prepare_numba_functions_on_the_fly, prepare_numba_functions_planned)

__all__ = ["FMM_Plan", "Get_Kernel_Functions", "Kernel_Apply", "Kernel_Form", "fmm_planner",
           "get_level_information", "get_print_function", "on_the_fly_fmm", "planned_fmm",
           "prepare_numba_functions_on_the_fly", "prepare_numba_functions_planned"]
