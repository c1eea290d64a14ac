"""MI355X-native (gfx950 HIP) lattice-Boltzmann hot path for MADDENING.

Layers:
  * libmaddening_lbm.so - HIP kernels behind the C ABI in include/lbm_c_api.h
  * maddening_amd.lbm.LBMKernel - PyTorch-ROCm front-end over that ABI
  * maddening_amd.node.LBMNode - drop-in for MADDENING's LBMNode
    (src/maddening/nodes/lbm.py:531-1036)
"""
from .lattice import LatticeDescriptor, d2q9, d3q19  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # torch-backed modules load on first use
    if name in ("LBMKernel",):
        from .lbm import LBMKernel
        return LBMKernel
    if name in ("LBMNode", "SimulationNode", "BoundaryInputSpec"):
        from . import node
        return getattr(node, name)
    raise AttributeError(name)
