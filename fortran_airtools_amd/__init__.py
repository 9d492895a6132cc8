"""FAIR tools reconstruction hot path on AMD Instinct MI355X (gfx950).

The product is libfair_hip.so (HIP kernels + C ABI, include/fair_hip.h) with two host mirrors
of the reference's module API: Fortran modules (fortran/) and these Python modules:

  sparse_matrices    CSRMatrix with matvecmul / transvecmul / row_norms / transpose / ...
  system_generation  Data, fanlineartomo, paralleltomo, linspace, add_white_noise
  airms              Opts, sirt, gradient_descent, normest, art

Importing the package does not touch the GPU; the library is loaded on first use and raises
FairError if it is missing.
"""
from ._lib import FairError  # noqa: F401

__all__ = ["FairError"]
