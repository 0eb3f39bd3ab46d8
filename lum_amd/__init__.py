"""lum_amd - MI355X-native voxel lighting hot path of Lum (radiance probe
update, map.comp voxelization, glossy reflection march) behind a C ABI
(include/lumx.h, include/clum.h).  liblumx.so holds the HIP kernels and the
host runtime; this package binds it for Python callers and tests."""
from ._lib import LumxError, load, RADIANCE_F32, RADIANCE_F16, RADIANCE_BF16  # noqa: F401

__all__ = ["LumxError", "load", "RADIANCE_F32", "RADIANCE_F16", "RADIANCE_BF16"]
