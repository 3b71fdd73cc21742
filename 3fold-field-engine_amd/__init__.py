"""MI355X-native block-sparse SDF build path: classify, compact, tile eval,
leaf-box pack, tiled CSG, the exact integer EDT and its mesh-stage caller, as hand-written HIP for
gfx950 behind a C ABI (include/fe_hip.h) and the reference's Python entry
points.  Importing the package loads libfe_hip.so and fails if it is absent.
"""
from . import _lib
from ._lib import LIB_PATH, NativeLibraryMissing, exported_symbols

_lib.lib()  # fail loudly at import when the HIP library is missing

from .edt import (EDTBatch, NativeEDT, NativeEDTLarge, NativeEDTVolume, PersistentPairEDT,  # noqa: E402
                  edt_device, signed_distance, signed_from_pair_device, squared_distance)

from . import recipes, field, batch, server, blocks, csg, octree, mesh, mesh_service, ikarus, ikarus_queries, multires  # noqa: E402
from .mesh import PreparedMeshField, surface_raster_and_flood_cuda  # noqa: E402
from .blocks import NODE, AdaptiveSampleBlocks, NativeAdaptiveBlocks  # noqa: E402
from .csg import evaluate, compose, gather_tile, NativeTileReader  # noqa: E402
from .field import (GridMeta, SparseField, FieldPlan, klassificera_och_evaluera_testfalt,  # noqa: E402
                    klassificera_och_evaluera_recept, klassificera_evaluera_rekonstruera_gpu_resident,
                    klassificera_evaluera_rekonstruera_recept, rekonstruera_tat, rekonstruera_tat_gpu,
                    reconstruct_solid_from_sparse, volym_via_gpu_reduktion, determinism_cert, konservering_check,
                    klassificera_och_evaluera_fran_tatt_falt)

__all__ = ["recipes", "field", "batch", "server", "blocks", "csg", "octree", "mesh", "mesh_service", "ikarus", "ikarus_queries", "multires",
           "PreparedMeshField",
           "surface_raster_and_flood_cuda", "NODE", "AdaptiveSampleBlocks",
           "NativeAdaptiveBlocks", "evaluate", "compose", "gather_tile", "NativeTileReader", "GridMeta", "SparseField", "FieldPlan", "klassificera_och_evaluera_testfalt",
           "klassificera_och_evaluera_recept", "klassificera_evaluera_rekonstruera_gpu_resident",
           "klassificera_evaluera_rekonstruera_recept", "rekonstruera_tat", "rekonstruera_tat_gpu",
           "reconstruct_solid_from_sparse", "volym_via_gpu_reduktion", "determinism_cert", "konservering_check",
           "klassificera_och_evaluera_fran_tatt_falt",
           "LIB_PATH", "NativeLibraryMissing", "exported_symbols", "NativeEDT", "NativeEDTLarge",
           "NativeEDTVolume", "PersistentPairEDT", "EDTBatch", "edt_device", "signed_distance", "signed_from_pair_device",
           "squared_distance"]
