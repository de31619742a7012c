"""flowcraft_amd — an MI355X-native Lattice Boltzmann engine for FlowCraft's LbmSolver step.

Hand-written HIP kernels for gfx950 behind a C ABI (include/flowcraft/fc_lbm.h,
libfc_lbm.so), with a Python mirror of the reference's LbmSolver class.
"""
from .lbm import LatticeModel, Layout, LbmSolver, Params, Stats, Storage, VoxelGrid  # noqa: F401

__all__ = ["LatticeModel", "Layout", "LbmSolver", "Params", "Stats", "Storage", "VoxelGrid"]
