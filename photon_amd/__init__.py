"""photon_amd — an MI355X-native rendering core for the Photon path tracer.

Host scene producers (binned-SAH BVH, light tables) in C++ and hand-written CDNA4 HIP kernels (wave64
BVH traversal, wavefront NEE+MIS integrator, SVGF) behind a C ABI:
  include/photon_types.h  shared plain-data layouts
  include/photon_host.h   host producers      -> photon_amd.host
  include/photon_hip.h    device library      -> photon_amd.gpu
"""
from . import types  # noqa: F401
from ._lib import NativeLibraryMissing  # noqa: F401

__all__ = ["types", "host", "gpu", "NativeLibraryMissing"]
