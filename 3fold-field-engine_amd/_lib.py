"""ctypes binding of libfe_hip.so, the gfx950 library behind every entry point.

The library must be built (build.py) before import; a missing or unloadable
library raises ImportError, so nothing in this package can silently run a
CPU substitute.
"""
import ctypes as ct
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfe_hip.so"

P = ct.c_void_p
I32, I64, F64, U32 = ct.c_int32, ct.c_int64, ct.c_double, ct.c_uint32

# name: (restype, argtypes); every symbol is declared in include/fe_hip.h.
SIGNATURES = {
    # EDT (reference-shaped)
    "fe_last_error": (ct.c_char_p, []),
    "edt_squared": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, P]),
    "edt_squared_large": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, P]),
    "edt_pair_create": (P, [ct.c_int, ct.c_int, ct.c_int]),
    "edt_pair_query": (ct.c_int, [P, P, F64, P, P]),
    "edt_pair_destroy": (None, [P]),
    # EDT (uncapped / device-resident)
    "fe_edt_squared_host": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, P]),
    "fe_edt_pair_create": (P, [ct.c_int, ct.c_int, ct.c_int]),
    "fe_edt_device": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, P, P]),
    "fe_edt_device_static": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, P]),
    "fe_edt_signed_device": (ct.c_int, [P, I64, F64, P, P]),
    "fe_edt_axis_pass": (ct.c_int, [P, P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P]),
    "fe_edt_batch_create": (P, [P, P, P, P, ct.c_int, ct.c_int, I64, I64, P]),
    "fe_edt_batch_run": (ct.c_int, [P, P, P, P, P]),
    "fe_edt_batch_destroy": (None, [P]),
    "fe_checksum_pair": (ct.c_int, [P, I64, F64, P, P, P]),
    "fe_octree_build": (P, [P, ct.c_int, ct.c_int, ct.c_int, P, P, P]),
    "fe_octree_info": (ct.c_int, [P, P, P, P]),
    "fe_octree_copy": (ct.c_int, [P, P, P, P, P, P]),
    "fe_octree_free": (None, [P]),
    "fe_octree_sample": (ct.c_int, [P, P, P, P, P, I64, F64, P, I64, P, P, P]),
    # field build
    "fe_field_build": (ct.c_int, [P, ct.c_int, ct.c_int, P, ct.c_int, P, ct.c_int, P, ct.c_int, ct.c_int, P]),
    "fe_field_build_counted": (ct.c_int, [P, ct.c_int, ct.c_int, P, ct.c_int, P, ct.c_int, P, ct.c_int, ct.c_int, P,
                                          P]),
    "fe_field_classify": (ct.c_int, [P, ct.c_int, ct.c_int, P, ct.c_int, ct.c_int, P]),
    "fe_field_tiles_ids": (ct.c_int, [P, P, I64, P]),
    "fe_field_reconstruct": (ct.c_int, [P, ct.c_int, I64, P, P, P]),
    "fe_field_checksum": (ct.c_int, [P, ct.c_int, F64, P, P]),
    "fe_dense_classify": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, P, P]),
    "fe_dense_tiles": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, I64, P, P]),
    "fe_multires_tiles": (ct.c_int, [P, P, ct.c_int, I64, P, P]),
    "fe_multires_band": (ct.c_int, [P, P, P, P, ct.c_int, P, P]),
    "fe_multires_query": (ct.c_int, [P, P, P, P, P, P, I64, P, P, P]),
    "fe_query_points": (ct.c_int, [P, P, I64, P, P]),
    "fe_copy_segments": (ct.c_int, [P, ct.c_int, P]),
    # leaf-box pack, lookups, CSG, tile reader
    "adaptive_pack": (P, [P, P, ct.c_int]),
    "adaptive_nodes": (I64, [P]),
    "adaptive_samples": (I64, [P]),
    "adaptive_copy": (ct.c_int, [P, P, I64, P, I64]),
    "adaptive_free": (None, [P]),
    "fe_pack_device": (P, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(ct.c_int), P]),
    "fe_pack_nodes": (I64, [P]),
    "fe_pack_samples": (I64, [P]),
    "fe_pack_copy": (ct.c_int, [P, P, I64, P, I64, P]),
    "fe_pack_free": (None, [P]),
    "fe_samples_at": (ct.c_int, [P, P, P, I64, P, P]),
    "fe_query_packed": (ct.c_int, [P, P, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(ct.c_double), F64, P, I64, P, P]),
    "fe_unpack": (ct.c_int, [P, I64, P, ct.c_int, ct.c_int, ct.c_int, P, P]),
    "csg_gather": (ct.c_int, [P, I64, P, I64, P, P, P]),
    "csg_execute": (ct.c_int, [P, P, P, P, I32, P, P, I32, P]),
    "fe_csg_check": (ct.c_int, [P, I64, P, I64, P, P, P]),
    "fe_csg_workspace_bytes": (I64, [ct.c_int32, P, P, ct.c_int32]),
    "fe_csg_run": (ct.c_int, [P, P, P, P, ct.c_int32, P, P, ct.c_int32, P, P, P, I64, P]),
    "fe_csg_device": (ct.c_int, [P, P, P, P, I32, P, P, I32, P, P]),
    # mesh stage (reference-shaped)
    "columns_create": (ct.c_int, [ct.c_char_p, P, U32, U32, U32, P, ct.POINTER(P)]),
    "columns_query": (ct.c_int, [P, P, U32, P, P, P]),
    "columns_query_compact": (ct.c_int, [P, P, U32, P, P, P, ct.POINTER(U32)]),
    "columns_destroy": (ct.c_int, [ct.POINTER(P)]),
    "column_mask_surface": (ct.c_int, [P, P, ct.c_int, P, P, ct.c_int, ct.c_int, ct.c_int, P, P]),
    "fused_create": (ct.c_int, [ct.c_char_p, P, U32, P, P, ct.c_int, ct.c_int, ct.c_int, F64, P, ct.POINTER(P)]),
    "fused_query": (ct.c_int, [P, P, P, P]),
    "fused_destroy": (ct.c_int, [ct.POINTER(P)]),
    # mesh stage (uncapped / device-resident)
    "fe_mesh_create": (ct.c_int, [P, U32, P, P, ct.c_int, ct.c_int, ct.c_int, F64, P, ct.c_int, ct.POINTER(P)]),
    "fe_gwn_create": (ct.c_int, [P, U32, P, ct.c_int, ct.c_int, ct.c_int, F64, ct.POINTER(P)]),
    "fe_mesh_query": (ct.c_int, [P, P, P, P]),
    "fe_mesh_destroy": (ct.c_int, [ct.POINTER(P)]),
    "fe_mesh_run": (ct.c_int, [P, P]),
    "fe_mesh_status": (ct.c_int, [P]),
    "fe_mesh_buffers": (ct.c_int, [P] + [ct.POINTER(P)] * 7 + [ct.POINTER(ct.c_int)]),
    "fe_gwn_device": (ct.c_int, [P, I64, P, I64, P, P, P, P]),
    # IKARUS trees
    "fe_ikarus_eval": (ct.c_int, [P, ct.c_int, P, P, P, P, I64, P, P]),
    "fe_ikarus_classify": (ct.c_int, [P, ct.c_int, P, P, P, P, P, ct.c_int, P, P, P]),
    "fe_ikarus_tiles": (ct.c_int, [P, ct.c_int, P, P, P, P, P, ct.c_int, P, I64, P, P]),
    "fe_ikarus_check": (ct.c_int, [P, ct.c_int, I64, ct.c_int]),
    "fe_ikarus_fault": (ct.c_int, [P]),
    "fe_ikarus_grid": (ct.c_int, [P, ct.c_int, P, P, P, P, F64, P, ct.c_float, ct.c_float, P, P, P, P]),
    "fe_ikarus_refine": (ct.c_int, [P, ct.c_int, P, P, P, P, F64, P, P, I64, ct.c_int, ct.c_float, P, P]),
    "fe_ikarus_pairs_eval": (ct.c_int, [P, ct.c_int, P, P, P, P, P, ct.c_int, I64, P, P]),
    "fe_ikarus_band_eval": (ct.c_int, [P, P, ct.c_int, P, P, P, ct.c_int, P, I64, P, P]),
}

_lib = None


class NativeLibraryMissing(ImportError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeLibraryMissing(f"{LIB_PATH} is not built; run build.py (hipcc, gfx950)")
        # One HIP runtime per process: torch's wheel carries its own
        # libamdhip64.so.7.  Loading torch first registers that soname, so the
        # library binds to the same runtime (and the same device context)
        # instead of starting a second HSA runtime that would find no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        try:
            L = ct.CDLL(str(LIB_PATH))
        except OSError as e:
            raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(SIGNATURES)


def stream_handle(stream=None):
    """Raw hipStream_t of a torch stream (None: torch's current stream)."""
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ct.c_void_p(s.cuda_stream)


def ptr(t):
    return ct.c_void_p(t.data_ptr())
