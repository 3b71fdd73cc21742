"""Mesh-stage test inputs: the reference's captured cases (tests/golden/mesh_stage.*),
closed and open meshes generated here, and the oracle's version of the whole
stage.  Test infrastructure only."""
import hashlib
import json
from pathlib import Path

import numpy as np

import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
GOLDEN_DIR = (0.3819660112501051, 0.6180339887498949, 0.2360679774997897)
JITTER = 1.0e-4


def digest(a):
    """sha256 of the array bytes (probes/field_engine/field_mesh_native_mask_stage_probe.py:20)."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fixtures():
    meta = json.loads((GOLDEN / "mesh_stage.json").read_text())
    with np.load(GOLDEN / "mesh_stage.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    return meta, arrays


# ------------------------------------------------------------------ grid
def window(v, pitch, lo, pad=2):
    """(gmin, shape) of FE/faltkarna_v1_mesh_to_sdf.py:153-167 without a global shape."""
    gmin = np.floor((v.min(0) - lo) / pitch).astype(np.int64) - pad
    gmax = np.ceil((v.max(0) - lo) / pitch).astype(np.int64) + pad + 1
    gmax = np.maximum(gmax, gmin)
    return gmin, tuple(int(x) for x in gmax - gmin)


def column_xy(anchor, shape, pitch):
    """One +z query per grid column (FE/mesh_field_cuda_columns_v1.py:56-59)."""
    ij = np.indices(shape[:2]).reshape(2, -1).T
    xy = anchor[:2] + (ij + 0.0) * pitch
    xy[:, 0] += pitch * JITTER * GOLDEN_DIR[0]
    xy[:, 1] += pitch * JITTER * GOLDEN_DIR[1]
    return np.ascontiguousarray(xy)


def layer_z(anchor, nz, pitch):
    """Sample z of each layer (FE/mesh_fused_handle_v1.py:23)."""
    return anchor[2] + (np.arange(nz, dtype=np.float64) + 0.0) * pitch


def sample_points(anchor, shape, pitch):
    """Every voxel sample relative to the anchor (FE/mesh_field_cuda_columns_v1.py:40-45)."""
    p = anchor + np.indices(shape).reshape(3, -1).T * pitch
    p[:, 0] += pitch * JITTER * GOLDEN_DIR[0]
    p[:, 1] += pitch * JITTER * GOLDEN_DIR[1]
    return p - anchor


def oracle_stage(v, f, pitch, lo, closed, cap=64):
    """The whole stage on the oracle: hits -> mask -> surface -> signed EDT."""
    v = np.asarray(v, np.float64)
    f = np.asarray(f, np.int64)
    gmin, shape = window(v, pitch, lo)
    anchor = lo + gmin.astype(np.float64) * pitch
    out = dict(gmin=gmin, shape=shape, anchor=anchor)
    if closed:
        counts, depths, signs = O.columns_scan(v[f], column_xy(anchor, shape, pitch), cap)
        solid, surface, flags = O.column_mask(counts, depths, signs, shape, layer_z(anchor, shape[2], pitch))
        out.update(counts=counts, depths=depths, signs=signs, flags=flags)
    else:
        w = O.gwn((v - anchor)[f], sample_points(anchor, shape, pitch))
        solid = (np.abs(w) > .5).reshape(shape)
        surface = O.surface_of(solid)
        out.update(w=w)
    out.update(solid=solid, surface=surface)
    if solid.any() and not solid.all():
        out["distance"] = O.edt_pair(solid, pitch)[0]
    return out


def reference_case(meta, arrays, name):
    """(vertices, faces, pitch, lo, closed) of a captured case with a mesh."""
    c = meta["cases"][name]
    return arrays[f"{name}_vertices"], arrays[f"{name}_faces"], c["pitch"], np.array(c["lo"]), c["closed"]


def capacity_case(meta, arrays):
    """32 stacked 4 mm boxes, 10 mm apart, and one query (field_cuda_columns_probe.py:77-84)."""
    cap = meta["capacity"]
    tri = arrays["capacity_vertices"][arrays["capacity_faces"]]
    corners = np.concatenate([tri + [0, 0, cap["step"] * i] for i in range(cap["stack"])])
    return corners, np.array([cap["xy"]])


def plate_columns(meta, arrays, p):
    """Per-column hits of a captured plate case, its shape and layer z."""
    name = f"plate_{p}"
    c = meta["cases"][name]
    shape = tuple(c["shape"])
    counts, depths, signs = O.hits_to_columns(arrays[f"{name}_columns"].astype(np.int64), arrays[f"{name}_depths"],
                                              arrays[f"{name}_signs"].astype(np.int32), shape[0] * shape[1])
    zs = c["anchor_z"] + (np.arange(shape[2], dtype=np.float64) + 0.0) * c["pitch"]
    return counts, depths, signs, shape, zs


# ------------------------------------------------------------------ generated meshes
def signed_volume(v, f):
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6)


def outward(v, f):
    return f if signed_volume(v, f) > 0 else f[:, ::-1].copy()


def star_sphere(n_lat, n_lon, radius, center, seed=None, wobble=0.0):
    """Closed, consistently wound UV sphere; with wobble, a star-shaped blob."""
    th = np.pi * np.arange(1, n_lat) / n_lat
    ph = 2 * np.pi * np.arange(n_lon) / n_lon
    T, Pn = np.meshgrid(th, ph, indexing="ij")
    r = np.full(T.shape, float(radius))
    if wobble:
        rng = np.random.default_rng(seed)
        for _ in range(4):
            k, m, a, s = rng.integers(1, 5), rng.integers(1, 5), rng.uniform(-wobble, wobble), rng.uniform(0, 6)
            r = r * (1 + a * np.sin(k * T + s) * np.cos(m * Pn + s))
    ring = np.stack([r * np.sin(T) * np.cos(Pn), r * np.sin(T) * np.sin(Pn), r * np.cos(T)], -1).reshape(-1, 3)
    v = np.concatenate([[[0, 0, radius]], ring, [[0, 0, -radius]]]) + np.asarray(center, np.float64)
    L, n = n_lat - 1, n_lon
    idx = lambda i, j: 1 + i * n + (j % n)  # noqa: E731
    faces = [(0, idx(0, j), idx(0, j + 1)) for j in range(n)]
    for i in range(L - 1):
        for j in range(n):
            a, b, c, d = idx(i, j), idx(i, j + 1), idx(i + 1, j), idx(i + 1, j + 1)
            faces += [(a, c, b), (b, c, d)]
    s = len(v) - 1
    faces += [(s, idx(L - 1, j + 1), idx(L - 1, j)) for j in range(n)]
    f = np.array(faces, np.int64)
    return v, outward(v, f)


def torus(R, r, nu, nv, center, tilt=0.0):
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    v = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1).reshape(-1, 3)
    ct, st = np.cos(tilt), np.sin(tilt)
    v = v @ np.array([[1, 0, 0], [0, ct, -st], [0, st, ct]]).T + np.asarray(center, np.float64)
    idx = lambda i, j: (i % nu) * nv + (j % nv)  # noqa: E731
    f = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i, j + 1), idx(i + 1, j + 1)
            f += [(a, b, d), (a, d, c)]
    f = np.array(f, np.int64)
    return v, outward(v, f)


def merge(*meshes):
    vs, fs, off = [], [], 0
    for v, f in meshes:
        vs.append(v)
        fs.append(f + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(fs)


def generated_cases():
    """(name, vertices, faces, pitch, lo) of closed meshes with uneven grids."""
    blob = star_sphere(24, 40, 9.0, (1.3, -0.7, 2.1), seed=5, wobble=0.18)
    tor = torus(11.0, 3.5, 48, 20, (0.4, 0.2, -0.3), tilt=0.6)
    pair = merge(star_sphere(16, 24, 5.0, (-9.0, 0.5, 0.0)), torus(6.0, 2.0, 40, 16, (8.5, -1.0, 1.0), tilt=1.1))
    out = []
    for name, (v, f), pitch in (("blob", blob, 0.7), ("torus", tor, 0.45), ("pair", pair, 0.6)):
        out.append((name, v, f, pitch, v.min(0) - np.array([1.7, 2.3, 1.1])))
    return out
