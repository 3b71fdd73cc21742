"""ctypes front end of the C oracle (oracle/fe_oracle.c).  TEST INFRASTRUCTURE.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module; the product path never does.  `lib()` builds the library with
make on first use if it is missing.
"""
import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libfe_oracle.so"
REF_DIR = HERE / "_ref"

NODE = np.dtype([("lo", "<i4", (3,)), ("shape", "<i4", (3,)), ("stored_shape", "<i4", (3,)),
                 ("children", "<i4", (2,)), ("offset", "<i8")])
PRIM = np.dtype([("kind", "<i4"), ("op", "<i4"), ("cx", "<f4"), ("cy", "<f4"), ("cz", "<f4"),
                 ("a", "<f4"), ("b", "<f4"), ("c", "<f4")])
GRID = np.dtype([("x0", "<f4"), ("y0", "<f4"), ("z0", "<f4"), ("pitch", "<f4"), ("margin", "<f4"),
                 ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("block", "<i4")])
INSTR = np.dtype([("op", "<i4"), ("left", "<i4"), ("right", "<i4"), ("pad", "<i4"), ("offset", "<f8")])

_lib = None
P = ct.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if Path("/root/reference").is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
        L = ct.CDLL(str(LIB_PATH))
        sig = {
            "or_edt_sentinel": (ct.c_int64, [ct.c_int] * 3),
            "or_edt_exhaustive": (None, [P, ct.c_int, ct.c_int, ct.c_int, P]),
            "or_edt_envelope": (None, [P, ct.c_int, ct.c_int, ct.c_int, P]),
            "or_edt_signed": (None, [P, P, ct.c_int64, ct.c_double, P]),
            "or_edt_pair": (None, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_double, P, P]),
            "or_prim_eval": (ct.c_float, [P, ct.c_float, ct.c_float, ct.c_float]),
            "or_recipe_eval_points": (None, [P, ct.c_int, P, ct.c_int64, P]),
            "or_classify": (None, [P, ct.c_int, P, P, P]),
            "or_compact": (ct.c_int64, [P, ct.c_int64, P]),
            "or_eval_tiles": (None, [P, ct.c_int, P, P, ct.c_int64, ct.c_int, P]),
            "or_reconstruct": (None, [P, P, ct.c_int64, P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P]),
            "or_checksum_i64": (ct.c_int64, [P, ct.c_int64, ct.c_double]),
            "or_classify_dense": (None, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, P]),
            "or_eval_tiles_dense": (None, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, ct.c_int64, P]),
            "or_pack": (ct.c_int, [P, P, ct.c_int, ct.POINTER(P), ct.POINTER(ct.c_int64), ct.POINTER(P),
                                   ct.POINTER(ct.c_int64)]),
            "or_free": (None, [P]),
            "or_gather": (ct.c_int, [P, ct.c_int64, P, ct.c_int64, P, P, P]),
            "or_csg_execute": (ct.c_int, [P, P, P, P, ct.c_int32, P, P, ct.c_int32, P]),
            "or_samples_at": (None, [P, P, P, ct.c_int64, P]),
            "or_query": (None, [P, P, P, P, ct.c_double, P, ct.c_int64, P]),
            "or_octree_build": (ct.c_int64, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_double, P, ct.c_int64, P,
                                             ct.c_int64, ct.POINTER(ct.c_int64), P]),
            "or_octree_sample": (ct.c_int, [P, P, P, P, P, ct.c_int64, ct.c_double, P, ct.c_int64, P]),
            "or_columns_scan": (None, [P, ct.c_int64, P, ct.c_int64, ct.c_int, P, P, P]),
            "or_column_mask": (ct.c_int, [P, P, P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, P, P]),
            "or_gwn": (ct.c_int, [P, ct.c_int64, P, ct.c_int64, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


# ---------------------------------------------------------------- EDT
def edt_exhaustive(binary):
    b = np.ascontiguousarray(binary, dtype=np.uint8)
    out = np.empty(b.shape, np.int64)
    lib().or_edt_exhaustive(_p(b), *b.shape, _p(out))
    return out


def edt_envelope(binary):
    b = np.ascontiguousarray(binary, dtype=np.uint8)
    out = np.empty(b.shape, np.int32)
    lib().or_edt_envelope(_p(b), *b.shape, _p(out))
    return out


def edt_pair(mask, pitch):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty(m.shape, np.float32)
    sq = np.empty((2,) + m.shape, np.int32)
    lib().or_edt_pair(_p(m), *m.shape, float(pitch), _p(out), _p(sq))
    return out, sq


def edt_signed_from_squares(sq0, sq1, pitch):
    a = np.ascontiguousarray(sq0, dtype=np.int32)
    b = np.ascontiguousarray(sq1, dtype=np.int32)
    out = np.empty(a.shape, np.float32)
    lib().or_edt_signed(_p(a), _p(b), a.size, float(pitch), _p(out))
    return out


# ---------------------------------------------------------------- field build
def grid_record(x0, y0, z0, pitch, nx, ny, nz, block, margin):
    g = np.zeros((), GRID)
    g["x0"], g["y0"], g["z0"], g["pitch"], g["margin"] = x0, y0, z0, pitch, margin
    g["nx"], g["ny"], g["nz"], g["block"] = nx, ny, nz, block
    return g


def recipe_eval_points(prims, pts):
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    out = np.empty(len(pts), np.float32)
    lib().or_recipe_eval_points(_p(prims), len(prims), _p(pts), len(pts), _p(out))
    return out


def build_sparse(prims, grid, corrupt=False):
    """classify -> ascending compaction -> tiles; returns (kind, cma, ids, tiles)."""
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    g = np.ascontiguousarray(grid, dtype=GRID)
    B = int(np.ravel(g["block"])[0])
    nb = [-(-int(np.ravel(g[k])[0]) // B) for k in ("nx", "ny", "nz")]
    n_blocks = nb[0] * nb[1] * nb[2]
    kind = np.empty(n_blocks, np.int32)
    cma = np.empty(n_blocks, np.float32)
    L = lib()
    L.or_classify(_p(prims), len(prims), _p(g), _p(kind), _p(cma))
    ids = np.empty(n_blocks, np.int32)
    n_active = L.or_compact(_p(kind), n_blocks, _p(ids))
    ids = ids[:n_active].copy()
    tiles = np.empty((n_active, B, B, B), np.float32)
    L.or_eval_tiles(_p(prims), len(prims), _p(g), _p(ids), n_active, 1 if corrupt else 0, _p(tiles))
    return kind, cma, ids, tiles


def reconstruct(kind, ids, tiles, nx, ny, nz, block):
    kind = np.ascontiguousarray(kind, np.int32)
    ids = np.ascontiguousarray(ids, np.int32)
    tiles = np.ascontiguousarray(tiles, np.float32)
    out = np.empty((nx, ny, nz), np.float32)
    lib().or_reconstruct(_p(kind), _p(ids), len(ids), _p(tiles), nx, ny, nz, block, _p(out))
    return out


def checksum_i64(vals, scale=1.0e6):
    v = np.ascontiguousarray(vals, np.float32).ravel()
    return int(lib().or_checksum_i64(_p(v), v.size, float(scale)))


def classify_dense(dense, block, margin):
    d = np.ascontiguousarray(dense, np.float32)
    nb = [-(-n // block) for n in d.shape]
    kind = np.empty(nb[0] * nb[1] * nb[2], np.int32)
    lib().or_classify_dense(_p(d), *d.shape, block, float(margin), _p(kind))
    return kind


def tiles_dense(dense, block, ids):
    d = np.ascontiguousarray(dense, np.float32)
    ids = np.ascontiguousarray(ids, np.int32)
    tiles = np.empty((len(ids), block, block, block), np.float32)
    lib().or_eval_tiles_dense(_p(d), *d.shape, block, _p(ids), len(ids), _p(tiles))
    return tiles


# ---------------------------------------------------------------- pack / CSG / query
def pack(samples, leaf_samples=512):
    a = np.ascontiguousarray(samples, np.float32)
    shape = np.array(a.shape, np.int32)
    pn, pp = P(), P()
    nn, npay = ct.c_int64(), ct.c_int64()
    L = lib()
    if L.or_pack(_p(a), _p(shape), int(leaf_samples), ct.byref(pn), ct.byref(nn), ct.byref(pp), ct.byref(npay)):
        raise ValueError("non-finite samples")
    nodes = np.empty(nn.value, NODE)
    payload = np.empty(npay.value, np.float32)
    ct.memmove(_p(nodes), pn.value, nodes.nbytes)
    ct.memmove(_p(payload), pp.value, payload.nbytes)
    L.or_free(pn)
    L.or_free(pp)
    return nodes, payload


def gather(nodes, payload, lower, upper):
    lo = np.ascontiguousarray(lower, np.int32)
    hi = np.ascontiguousarray(upper, np.int32)
    out = np.empty(tuple(np.maximum(hi - lo, 1)), np.float32)
    code = lib().or_gather(_p(nodes), len(nodes), _p(payload), len(payload), _p(lo), _p(hi), _p(out))
    return code, out


def csg_execute(fields, program, shape):
    """fields: list of (nodes, payload); program: INSTR array.  Returns (code, dense)."""
    nodes = (P * len(fields))(*[_p(f[0]) for f in fields])
    pays = (P * len(fields))(*[_p(f[1]) for f in fields])
    counts = np.array([len(f[0]) for f in fields], np.int64)
    lens = np.array([len(f[1]) for f in fields], np.int64)
    sh = np.array(shape, np.int32)
    prog = np.ascontiguousarray(program, INSTR)
    out = np.empty(tuple(shape), np.float32)
    code = lib().or_csg_execute(nodes, _p(counts), pays, _p(lens), len(fields), _p(sh), _p(prog), len(prog),
                                _p(out))
    return code, out


def samples_at(nodes, payload, idx):
    idx = np.ascontiguousarray(idx, np.int64)
    out = np.empty(len(idx), np.float32)
    lib().or_samples_at(_p(nodes), _p(payload), _p(idx), len(idx), _p(out))
    return out


def query(nodes, payload, shape, origin, pitch, points):
    pts = np.ascontiguousarray(points, np.float64)
    sh = np.array(shape, np.int32)
    org = np.ascontiguousarray(origin, np.float64)
    out = np.empty(len(pts), np.float32)
    lib().or_query(_p(nodes), _p(payload), _p(sh), _p(org), float(pitch), _p(pts), len(pts), _p(out))
    return out


def ref_lib(name):
    """A reference C++ native compiled by `make -C oracle ref` (None if absent)."""
    path = REF_DIR / f"libref_{name}.so"
    return ct.CDLL(str(path)) if path.exists() else None


# ---------------------------------------------------------------- 2-bit adaptive block tree
OCT_LOOKUP = np.array([[sum(((byte >> (2 * i)) & 3) == code for i in range(4)) for byte in range(256)]
                       for code in range(4)], dtype=np.int32)


def octree_build(dense, pitch):
    """FE/field_adaptive_blocks_v1.py:97-138: node codes and values from the C
    restatement, then the packing (four codes per byte, code i at bits 2i) and
    the rank checkpoints every 64 bytes restated in numpy.  Returns a dict with
    the reference's AdaptiveBlocks fields and the per-level code counts."""
    d = np.ascontiguousarray(dense, np.float32)
    nx, ny, nz = d.shape
    root_shape = tuple(-(-n // 8) for n in d.shape)
    cap = 9 * int(np.prod(root_shape)) * 8 + 64
    while True:
        codes = np.empty(cap, np.uint8)
        values = np.empty(max(d.size, 1), np.float32)
        nv = ct.c_int64(0)
        hist = np.zeros(16, np.int64)
        count = lib().or_octree_build(_p(d), nx, ny, nz, float(pitch), _p(codes), cap, _p(values), values.size,
                                      ct.byref(nv), _p(hist))
        if count >= 0:
            break
        cap *= 4
    codes = codes[:count]
    padded = np.pad(codes, (0, (-len(codes)) % 4))
    packed = np.bitwise_or.reduce(padded.reshape(-1, 4) << np.array([0, 2, 4, 6], np.uint8), axis=1).astype(np.uint8)
    prefixes = []
    for code in (2, 3):
        cumulative = np.concatenate(([0], np.cumsum(OCT_LOOKUP[code, packed], dtype=np.int32)))
        prefixes.append(cumulative[np.arange(0, len(packed) + 1, 64)].astype(np.int32))
    return dict(shape=d.shape, pitch=float(pitch), root_shape=root_shape, root_count=int(np.prod(root_shape)),
                node_count=int(count), codes=codes, packed=packed, branch_prefix=prefixes[0],
                value_prefix=prefixes[1], values=values[:nv.value].copy(), level_counts=hist.reshape(4, 4))


def octree_sample(tree, indices):
    """FE/field_adaptive_blocks_v1.py:52-82 through the C restatement."""
    ijk = np.ascontiguousarray(indices, np.int64)
    out = np.empty(len(ijk), np.float32)
    rs = np.ascontiguousarray(tree["root_shape"], np.int32)
    vals = tree["values"] if len(tree["values"]) else np.zeros(1, np.float32)
    status = lib().or_octree_sample(_p(tree["packed"]), _p(tree["branch_prefix"]), _p(tree["value_prefix"]),
                                    _p(vals), _p(rs), tree["root_count"], tree["pitch"], _p(ijk), len(ijk), _p(out))
    if status:
        raise RuntimeError("Invalid tree depth")
    return out


# ---------------------------------------------------------------- mesh stage
def columns_scan(corners, xy, cap=64):
    """Per-column hits of FE/cuda_columns_v1/columns.cu:28-48 (face order):
    (counts int32[q], depths f64[q, cap], signs int32[q, cap])."""
    c = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1, 9)
    q = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    counts = np.empty(len(q), np.int32)
    depths = np.zeros((len(q), cap), np.float64)
    signs = np.zeros((len(q), cap), np.int32)
    lib().or_columns_scan(_p(c), len(c), _p(q), len(q), cap, _p(counts), _p(depths), _p(signs))
    return counts, depths, signs


def sorted_hits(counts, depths, signs):
    """Flatten per-column hits and order them as the reference's handles do
    (FE/rt_columns_handle_v1.py:47-49): lexsort by column, depth, sign."""
    cap = depths.shape[1]
    valid = np.arange(cap)[None, :] < np.minimum(counts, cap)[:, None]
    c = np.broadcast_to(np.arange(len(counts))[:, None], valid.shape)[valid]
    z = depths[valid]
    s = signs[valid]
    order = np.lexsort((s, z, c))
    return c[order], z[order], s[order]


def column_mask(counts, depths, signs, shape, zs):
    """(solid bool, surface bool, flags) of the fused mask chain."""
    nx, ny, nz = shape
    cap = depths.shape[1]
    cn = np.ascontiguousarray(counts, dtype=np.int32)
    d = np.ascontiguousarray(depths, dtype=np.float64)
    s = np.ascontiguousarray(signs, dtype=np.int32)
    z = np.ascontiguousarray(zs, dtype=np.float64)
    mask = np.empty(shape, np.uint8)
    surface = np.empty(shape, np.uint8)
    flags = lib().or_column_mask(_p(cn), _p(d), _p(s), cap, nx, ny, nz, _p(z), _p(mask), _p(surface))
    return mask.view(np.bool_), surface.view(np.bool_), flags


def hits_to_columns(c, z, s, n_columns, cap=64):
    """Inverse of sorted_hits: per-column arrays from flat (column, depth, sign)."""
    counts = np.bincount(np.asarray(c, np.int64), minlength=n_columns).astype(np.int32)
    depths = np.zeros((n_columns, cap), np.float64)
    signs = np.zeros((n_columns, cap), np.int32)
    order = np.argsort(c, kind="stable")
    c, z, s = np.asarray(c)[order], np.asarray(z)[order], np.asarray(s)[order]
    start = np.r_[0, np.cumsum(counts)]
    k = np.arange(len(c)) - start[c]
    depths[c, k] = z
    signs[c, k] = s
    return counts, depths, signs


def gwn(corners, points):
    """Ordered generalized winding numbers (FE/generalized_winding_ordered_v1.py:11-32)."""
    c = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1, 9)
    q = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    w = np.empty(len(q), np.float64)
    if lib().or_gwn(_p(c), len(c), _p(q), len(q), _p(w)):
        raise ValueError("query coincides with a vertex")
    return w


def surface_of(solid):
    """Solid voxels on the grid border or with a 6-neighbour outside
    (FE/mesh_fused_native_v1/field.cu:31-36)."""
    m = np.asarray(solid, bool)
    inner = np.zeros_like(m)
    inner[1:-1, 1:-1, 1:-1] = (m[1:-1, 1:-1, 1:-1] & m[:-2, 1:-1, 1:-1] & m[2:, 1:-1, 1:-1] & m[1:-1, :-2, 1:-1] &
                               m[1:-1, 2:, 1:-1] & m[1:-1, 1:-1, :-2] & m[1:-1, 1:-1, 2:])
    return m & ~inner
