"""helfem_amd: MI355X-native Fock-build engine for HelFEM.

The shipped path is libhfgpu.so (hand-written gfx950 HIP kernels behind the
C ABI of include/hfgpu.h); this package is its Python host mirror.
"""
from ._lib import HfError, lib  # noqa: F401
from .atomic import AtomicBasis, AtomicBatch, Context, angular_basis, element_grid  # noqa: F401
from .diatomic import DiatomicBasis, diatomic_shells, mu_max  # noqa: F401
from .dftgrid import DFTGrid, parse_method  # noqa: F401

__all__ = ["AtomicBasis", "AtomicBatch", "Context", "DFTGrid", "DiatomicBasis", "HfError", "angular_basis", "diatomic_shells", "element_grid",
           "lib", "mu_max", "parse_method"]
