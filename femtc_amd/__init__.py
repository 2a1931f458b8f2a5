"""femtc_amd -- MI355X (gfx950) sum-factorized mass / stiffness operators.

Host-side mirror of the reference's operator interface over the C ABI of
libfemtc.so (include/femtc.h).  Importing the package does not load the
library; the first call does, and fails loudly if it is not built.
"""
from ._lib import (BF16, BF16_ROUNDED, F32, F64, HIGH_DEGREE_AUTO,
                   HIGH_DEGREE_GENERAL, MODE_CORRECT, MODE_REFERENCE, OP_KPM,
                   OP_MASS, OP_STIFFNESS, SCATTER_ASSEMBLED, SCATTER_AUTO,
                   SCATTER_DIRECT, SCATTER_ZRUN, FemtcError, lib)
from .linalg import Order, string2order
from .mesh import (basis_tables, basis_tables_weights, dofmap_structured,
                   even_odd_tables, gauss_rule, geometry_affine,
                   geometry_cell, gll_points)
from .operators import (apply_kernel, fill_zero, global_apply,
                        global_apply_affine, high_degree_kernel, kpm_apply,
                        mass_apply,
                        mass_operator, set_high_degree_kernel,
                        stiffness_apply, stiffness_operator)
from .mfma import global_apply_mfma, mfma_tables, mfma_tables_bytes
from .batch import Batch
from .plan import Plan
from .server import ResidentProblem, Server, latency_stats
from .solver import PCGSolver, global_diagonal

__all__ = [
    "Batch", "Plan", "PCGSolver", "ResidentProblem", "Server",
    "global_diagonal", "latency_stats",
    "BF16", "BF16_ROUNDED", "F32", "F64", "MODE_CORRECT", "MODE_REFERENCE", "OP_KPM", "OP_MASS",
    "OP_STIFFNESS", "SCATTER_ASSEMBLED", "SCATTER_AUTO", "SCATTER_DIRECT",
    "SCATTER_ZRUN", "HIGH_DEGREE_AUTO", "HIGH_DEGREE_GENERAL",
    "high_degree_kernel", "set_high_degree_kernel", "apply_kernel",
    "FemtcError", "lib", "Order", "string2order", "fill_zero",
    "basis_tables", "dofmap_structured", "gauss_rule", "geometry_affine",
    "gll_points", "global_apply", "kpm_apply", "mass_apply", "mass_operator",
    "stiffness_apply", "stiffness_operator", "global_apply_mfma",
    "mfma_tables", "mfma_tables_bytes",
]
