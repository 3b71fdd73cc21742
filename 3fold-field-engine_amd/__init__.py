"""MI355X-native block-sparse SDF build path: classify, compact, tile eval,
leaf-box pack, tiled CSG and the exact integer EDT, as hand-written HIP for
gfx950 behind a C ABI (include/fe_hip.h) and the reference's Python entry
points.  Importing the package loads libfe_hip.so and fails if it is absent.
"""
from . import _lib
from ._lib import LIB_PATH, NativeLibraryMissing, exported_symbols

_lib.lib()  # fail loudly at import when the HIP library is missing

from .edt import (NativeEDT, NativeEDTLarge, NativeEDTVolume, PersistentPairEDT,  # noqa: E402
                  edt_device, signed_distance, signed_from_pair_device, squared_distance)

__all__ = ["LIB_PATH", "NativeLibraryMissing", "exported_symbols", "NativeEDT", "NativeEDTLarge",
           "NativeEDTVolume", "PersistentPairEDT", "edt_device", "signed_distance", "signed_from_pair_device",
           "squared_distance"]
