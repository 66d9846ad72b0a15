"""MI355X-native CSR5 SpMV and conjugate gradient (gfx950 HIP kernels behind a C ABI).

Mirrors the CSR5 surface of dim (rurabori/diploma_research): Csr5.from_csr,
Csr5.spmv, conjugate_gradient. Importing this package does not load the GPU
library; the first call does (and fails loudly if it was not built).
"""
from .csr5 import (Calibrator, CgResult, Coo, Csr, Csr5, Dimensions, SpmvData, conjugate_gradient,
                   conjugate_gradient_multi)
from .native import DimError

__all__ = ["Calibrator", "CgResult", "Coo", "Csr", "Csr5", "Dimensions", "SpmvData", "conjugate_gradient",
           "conjugate_gradient_multi", "DimError"]
