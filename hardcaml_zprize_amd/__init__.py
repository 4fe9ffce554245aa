"""MI355X-native Goldilocks NTT and BLS12-377 G1 MSM (host side over the C ABI).

The compute runs in hardcaml_zprize_amd/libzpgpu.so (HIP, gfx950); this
package only mirrors the reference's host-driver interfaces on top of it.
"""
from ._lib import LIB_PATH, ZpError, lib, last_error  # noqa: F401
