"""IKARUS expression trees on the device.

A part is an expression tree of plain dict nodes (the format of
FE/ikarus_v1/expr.py: JSON round-trips it, `save`/`load` wrap the versioned
envelope).  Leaves: halfspace, sphere, box, cylinder, cone, rect_frustum,
gridfield (trilinear samples); interior nodes: transform (affine 4x4),
axis_project, offset, scale, and union / intersect / subtract with a smooth
blend radius k (k = 0: hard).  Sign convention: < 0 inside.

`compile_tree` turns a tree into a program for the device interpreter in
csrc/ikarus.hip: 16-byte instructions (opcode, argument, parameter offset) in
post order and an fp32 parameter vector.  Parameters are rounded to fp32
from the double values the reference bakes into its generated kernels
(FE/ikarus_v1/eval_warp.py:160-320; rect_frustum as
FE/ikarus_v1/eval_warp_uniform_v2.py:180-214).  Each leaf is preceded by its
ancestors' point maps, applied from the world point outward-in; the two
operands of a boolean are evaluated larger-stack-need first (the result does
not depend on the order), so the LDS value stack stays shallow.

Entry points: `eval_batch(node, points)` and `gradient_batch(node, points, h)`
(FE/ikarus_v1/eval_warp.py:471-485, 718-735), and
`klassificera_och_evaluera_ikarus(meta, node)`: the block-sparse build of a
field from a tree (eight-corner classification, host compaction, tiles), in
the structure of the plain per-call build.
"""
import json
import math

import numpy as np

from . import _lib

IKARUS_VERSION = "1.0"

# opcodes of csrc/ikarus.hip
WORLD, XFORM, PROJ, SCALE_IN = 0, 1, 2, 3
HALFSPACE, SPHERE, BOX, CYL, CONE, FRUSTUM, GRID = 10, 11, 12, 13, 14, 15, 16
OFFSET, SCALE_OUT, UNION, INTERSECT, SUBTRACT = 20, 21, 22, 23, 24
STACK_SLOTS = 16

GRID_REC = np.dtype([("off", "<i8"), ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("is_sdf", "<i4"),
                     ("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("pitch", "<f4"), ("iso", "<f4"), ("pad", "<f4")])
assert GRID_REC.itemsize == 48
AXES = {"x": 0, "y": 1, "z": 2}


# ------------------------------------------------------------------ 4x4 affine maps
# Transform nodes carry a row-major 4x4 affine matrix as nested lists (the
# JSON form of FE/ikarus_v1/expr.py:24-83).
def mat4_identity():
    return np.eye(4).tolist()


def mat4_translate(tx, ty, tz):
    m = np.eye(4)
    m[:3, 3] = (tx, ty, tz)
    return m.tolist()


def mat4_rot_axis(axis, deg):
    """Rotation by `deg` degrees about coordinate axis 'x', 'y' or 'z'
    (right-handed: the plane (axis+1, axis+2) turns counter-clockwise)."""
    if axis not in AXES:
        raise ValueError(f"rotation axis must be one of x, y, z (got {axis!r})")
    k = AXES[axis]
    i, j = (k + 1) % 3, (k + 2) % 3
    rad = math.radians(deg)
    m = np.eye(4)
    m[i, i] = m[j, j] = math.cos(rad)
    m[i, j], m[j, i] = -math.sin(rad), math.sin(rad)
    return m.tolist()


def mat4_mul(a, b):
    return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).tolist()


def mat4_inverse_affine(m):
    """World -> local map of an affine 4x4: the 3x3 block inverted through its
    adjugate, whose column i is the cross product of rows i+1 and i+2, then
    the translation.  Each entry is one double product difference divided by
    the determinant, and every sum runs left to right, so the values equal
    those of the cofactor expansion in FE/ikarus_v1/expr.py:57-83 bit for bit
    (the fp32 parameters of a transform then round alike)."""
    rows = [np.asarray(m[r][:3], np.float64) for r in range(3)]
    adj = [np.cross(rows[(c + 1) % 3], rows[(c + 2) % 3]) for c in range(3)]
    det = rows[0][0] * adj[0][0] + rows[0][1] * adj[0][1] + rows[0][2] * adj[0][2]
    if not abs(det) >= 1e-15:
        raise ValueError("transform matrix is singular")
    inv = np.empty((3, 3))
    for c in range(3):
        inv[:, c] = adj[c] / det
    t = [float(m[r][3]) for r in range(3)]
    out = np.eye(4)
    out[:3, :3] = inv
    for r in range(3):
        out[r, 3] = -(inv[r, 0] * t[0] + inv[r, 1] * t[1] + inv[r, 2] * t[2])
    return out.tolist()


# ------------------------------------------------------------------ nodes (FE/ikarus_v1/expr.py:90-211)
def halfspace(normal, offset):
    nx, ny, nz = normal
    nrm = math.sqrt(nx * nx + ny * ny + nz * nz)
    if not nrm > 1e-12:
        raise ValueError("halfspace normal must be nonzero")
    return {"op": "halfspace", "normal": [nx / nrm, ny / nrm, nz / nrm], "offset": float(offset)}


def sphere(radius, center=(0.0, 0.0, 0.0)):
    return {"op": "sphere", "radius": float(radius), "center": list(map(float, center))}


def box(half_extents, center=(0.0, 0.0, 0.0)):
    return {"op": "box", "half_extents": list(map(float, half_extents)), "center": list(map(float, center))}


def cylinder(radius, height, axis="z", center=(0.0, 0.0, 0.0)):
    return {"op": "cylinder", "radius": float(radius), "height": float(height), "axis": axis,
            "center": list(map(float, center))}


def cone(radius1, radius2, height, axis="z", center=(0.0, 0.0, 0.0)):
    return {"op": "cone", "radius1": float(radius1), "radius2": float(radius2), "height": float(height),
            "axis": axis, "center": list(map(float, center))}


def rect_frustum(z0, z1, hx0, hy0, hx1, hy1, cx0=0.0, cy0=0.0, cx1=0.0, cy1=0.0):
    if not z1 > z0:
        raise ValueError("rect_frustum requires z1 > z0")
    return {"op": "rect_frustum", "z0": float(z0), "z1": float(z1), "hx0": float(hx0), "hy0": float(hy0),
            "hx1": float(hx1), "hy1": float(hy1), "cx0": float(cx0), "cy0": float(cy0), "cx1": float(cx1),
            "cy1": float(cy1)}


def gridfield(npz_path, key="occupancy", mode="occupancy", iso=0.5, pitch_override=None):
    """A trilinear grid leaf read from an .npz (keys `key`, optional
    `global_origin_mm` (3,) and `pitch_mm`); mode "sdf" uses the values as
    distances, "occupancy" reads (iso - value) * pitch."""
    return {"op": "gridfield", "npz_path": npz_path, "key": key, "mode": mode, "iso": float(iso),
            "pitch_override": pitch_override}


def grid_leaf(values, origin=(0.0, 0.0, 0.0), pitch=1.0, mode="sdf", iso=0.5):
    """A grid leaf holding its samples in memory (not JSON-serialisable)."""
    return {"op": "gridfield", "values": np.asarray(values, np.float32), "origin": list(map(float, origin)),
            "pitch": float(pitch), "mode": mode, "iso": float(iso)}


def transform(child, mat4):
    return {"op": "transform", "child": child, "mat4": mat4}


def translate(child, tx, ty, tz):
    return transform(child, mat4_translate(tx, ty, tz))


def rotate(child, axis, deg, about=(0.0, 0.0, 0.0)):
    ax, ay, az = about
    m = mat4_mul(mat4_translate(ax, ay, az), mat4_mul(mat4_rot_axis(axis, deg), mat4_translate(-ax, -ay, -az)))
    return transform(child, m)


def axis_project(child, axis, ref):
    if axis not in (0, 1, 2):
        raise ValueError(f"axis_project: axis must be 0/1/2 (x/y/z), got {axis!r}")
    return {"op": "axis_project", "child": child, "axis": int(axis), "ref": float(ref)}


def offset(child, delta):
    return {"op": "offset", "child": child, "delta": float(delta)}


def scale(child, s):
    return {"op": "scale", "child": child, "s": float(s)}


def union(a, b, k=0.0):
    return {"op": "union", "a": a, "b": b, "k": float(k)}


def intersect(a, b, k=0.0):
    return {"op": "intersect", "a": a, "b": b, "k": float(k)}


def subtract(a, b, k=0.0):
    return {"op": "subtract", "a": a, "b": b, "k": float(k)}


def save(node, path):
    with open(path, "w") as f:
        json.dump({"ikarus_version": IKARUS_VERSION, "tree": node}, f, indent=1)
    return path


def load(path):
    with open(path) as f:
        data = json.load(f)
    if isinstance(data, dict) and "tree" in data and "ikarus_version" in data:
        return data["tree"]
    return data


def node_count(node):
    if not isinstance(node, dict):
        return 0
    return 1 + sum(node_count(node[k]) for k in ("child", "a", "b") if k in node)


def leaf_ops(node):
    out = set()

    def walk(n):
        if n["op"] in ("halfspace", "sphere", "box", "cylinder", "cone", "rect_frustum", "gridfield"):
            out.add(n["op"])
        for k in ("child", "a", "b"):
            if k in n:
                walk(n[k])
    walk(node)
    return out


# ------------------------------------------------------------------ compiler
def _grid_samples(node):
    """(fp32 samples, origin, pitch, iso, is_sdf) of a grid leaf
    (FE/ikarus_v1/eval_warp.py:433-450 for the .npz form)."""
    if "values" in node:
        arr = np.ascontiguousarray(node["values"], np.float32)
        return arr, node["origin"], float(node["pitch"]), float(node["iso"]), node["mode"] == "sdf"
    with np.load(node["npz_path"], allow_pickle=False) as d:
        arr = np.ascontiguousarray(d[node["key"]], np.float32)
        origin = d["global_origin_mm"] if "global_origin_mm" in d else np.zeros(3)
        pitch = node["pitch_override"] or float(d["pitch_mm"]) if "pitch_mm" in d else 1.0
    return arr, [float(x) for x in origin], float(pitch), float(node["iso"]), node["mode"] == "sdf"


def _grid_record(node, off):
    """(48-byte grid record as a tuple, fp32 samples) of a grid leaf whose
    samples start at element `off` of the grid data."""
    arr, origin, pitch, iso, is_sdf = _grid_samples(node)
    if arr.ndim != 3 or min(arr.shape) < 1:
        raise ValueError("grid leaf needs a 3-D array")
    return (off, *arr.shape, 1 if is_sdf else 0, *(np.float32(x) for x in origin), np.float32(pitch),
            np.float32(iso), 0.0), arr


def _need(node, memo):
    """Value-stack slots a subtree needs (Sethi-Ullman, the register top counted)."""
    key = id(node)
    if key in memo:
        return memo[key]
    op = node["op"]
    if op in ("union", "intersect", "subtract"):
        na, nb = _need(node["a"], memo), _need(node["b"], memo)
        r = max(na, nb) if na != nb else na + 1
    elif "child" in node:
        r = _need(node["child"], memo)
    else:
        r = 1
    memo[key] = r
    return r


class Program:
    """A compiled tree: instructions (int32 [n, 4]), fp32 parameters, grid
    records and samples, and the value-stack depth it needs."""

    def __init__(self, ins, params, grids, gdata, depth):
        self.ins = ins
        self.params = params
        self.grids = grids
        self.gdata = gdata
        self.depth = depth
        self._dev = {}

    def device(self, device):
        """Device copies (torch tensors), made once per device."""
        import torch
        dev = torch.device(device)
        if dev not in self._dev:
            def up(a, dtype):
                return torch.from_numpy(np.ascontiguousarray(a)).to(dev) if len(a) else torch.zeros(1, dtype=dtype,
                                                                                                  device=dev)
            self._dev[dev] = (up(self.ins.reshape(-1), torch.int32), up(self.params, torch.float32),
                              up(self.grids.view(np.uint8), torch.uint8), up(self.gdata, torch.float32))
        return self._dev[dev]


def compile_tree(node):
    ins, params, grids, chunks = [], [], [], []
    goff = [0]
    state = {"chain": None}

    def slot(values):
        off = len(params)
        params.extend(np.float32(v) for v in values)
        return off

    def emit(op, a=0, values=()):
        ins.append((op, a, slot(values), 0))

    def leaf(chain, op, a, values):
        if state["chain"] != [id(c[0]) for c in chain]:
            emit(WORLD)
            for c in chain:
                emit(*c[1])
            state["chain"] = [id(c[0]) for c in chain]
        emit(op, a, values)

    memo = {}

    def walk(n, chain):
        op = n["op"]
        if op == "halfspace":
            leaf(chain, HALFSPACE, 0, list(n["normal"]) + [n["offset"]])
        elif op == "sphere":
            leaf(chain, SPHERE, 0, list(n["center"]) + [n["radius"]])
        elif op == "box":
            leaf(chain, BOX, 0, list(n["center"]) + list(n["half_extents"]))
        elif op == "cylinder":
            leaf(chain, CYL, AXES[n["axis"]], list(n["center"]) + [n["radius"], n["height"] / 2.0])
        elif op == "cone":
            leaf(chain, CONE, AXES[n["axis"]], list(n["center"]) + [n["height"] / 2.0, n["height"], n["radius1"],
                                                                     n["radius2"]])
        elif op == "rect_frustum":
            z0, z1 = n["z0"], n["z1"]
            kx = (n["hx1"] - n["hx0"]) / (z1 - z0)
            ky = (n["hy1"] - n["hy0"]) / (z1 - z0)
            kcx = (n["cx1"] - n["cx0"]) / (z1 - z0)
            kcy = (n["cy1"] - n["cy0"]) / (z1 - z0)
            leaf(chain, FRUSTUM, 0, [z0, z1,
                                     -(kx + kcx), n["hx0"] + n["cx0"] - (kx + kcx) * z0,
                                     -(kx - kcx), n["hx0"] - n["cx0"] - (kx - kcx) * z0,
                                     -(ky + kcy), n["hy0"] + n["cy0"] - (ky + kcy) * z0,
                                     -(ky - kcy), n["hy0"] - n["cy0"] - (ky - kcy) * z0])
        elif op == "gridfield":
            rec, arr = _grid_record(n, goff[0])
            grids.append(rec)
            chunks.append(arr.reshape(-1))
            goff[0] += arr.size
            leaf(chain, GRID, len(grids) - 1, ())
        elif op == "transform":
            inv = mat4_inverse_affine(n["mat4"])
            walk(n["child"], chain + [(n, (XFORM, 0, [inv[r][c] for r in range(3) for c in range(4)]))])
        elif op == "axis_project":
            walk(n["child"], chain + [(n, (PROJ, int(n["axis"]), [n["ref"]]))])
        elif op == "scale":
            walk(n["child"], chain + [(n, (SCALE_IN, 0, [n["s"]]))])
            emit(SCALE_OUT, 0, [n["s"]])
        elif op == "offset":
            walk(n["child"], chain)
            emit(OFFSET, 0, [n["delta"]])
        elif op in ("union", "intersect", "subtract"):
            na, nb = _need(n["a"], memo), _need(n["b"], memo)
            if nb > na:  # evaluate b first: the top then holds a
                walk(n["b"], chain)
                walk(n["a"], chain)
                swapped = 1
            else:
                walk(n["a"], chain)
                walk(n["b"], chain)
                swapped = 0
            emit({"union": UNION, "intersect": INTERSECT, "subtract": SUBTRACT}[op], swapped, [n["k"]])
        else:
            raise ValueError(f"unknown op {op!r}")

    depth = _need(node, memo)
    if depth > STACK_SLOTS + 1:
        raise ValueError(f"tree needs {depth} value slots; the device stack holds {STACK_SLOTS + 1}")
    walk(node, [])
    ins_a = np.array(ins, np.int32).reshape(-1, 4)
    par = np.array(params, np.float32) if params else np.zeros(1, np.float32)
    grid_a = np.array(grids, GRID_REC) if grids else np.zeros(0, GRID_REC)
    gdata = np.concatenate(chunks).astype(np.float32) if chunks else np.zeros(0, np.float32)
    check_program(ins_a, par, grid_a)
    return Program(ins_a, par, grid_a, gdata, depth)


def check_program(ins, params, grids):
    """The library's host check of a compiled program (opcodes, parameter
    and grid ranges, stack depth); raises ValueError with its message."""
    ins = np.ascontiguousarray(ins, np.int32).reshape(-1, 4)
    L = _lib.lib()
    if L.fe_ikarus_check(ins.ctypes.data, len(ins), len(params), len(grids)):
        raise ValueError(L.fe_last_error().decode())


# ------------------------------------------------------------------ device entry points
def _program(node):
    return node if isinstance(node, Program) else compile_tree(node)


def _check(status, what):
    if status:
        raise RuntimeError(f"{what} failed with status {status}")


def eval_device(prog, points, out, stream=None):
    """Values of a compiled program at device points (n, 3) fp32 into out (n,) fp32."""
    ins, par, grids, gdata = prog.device(points.device)
    L = _lib.lib()
    P = _lib.ptr
    _check(L.fe_ikarus_eval(P(ins), len(prog.ins), P(par), P(grids), P(gdata), P(points), points.shape[0], P(out),
                            _lib.stream_handle(stream)), "ikarus eval")


def eval_batch(node, points_mm, device="cuda:0"):
    """(N, 3) points (mm) -> (N,) fp32 values (FE/ikarus_v1/eval_warp.py:471-485)."""
    import torch
    prog = _program(node)
    pts = torch.from_numpy(np.ascontiguousarray(np.asarray(points_mm).astype(np.float32)).reshape(-1, 3)).to(device)
    out = torch.empty(pts.shape[0], dtype=torch.float32, device=pts.device)
    eval_device(prog, pts, out)
    return out.cpu().numpy()


def gradient_batch(node, points_mm, h=0.05, device="cuda:0"):
    """Central differences at +-h per axis in one evaluation of 6N points
    (FE/ikarus_v1/eval_warp.py:718-735); (N, 3)."""
    points_mm = np.asarray(points_mm)
    n = points_mm.shape[0]
    offsets = np.array([[h, 0, 0], [-h, 0, 0], [0, h, 0], [0, -h, 0], [0, 0, h], [0, 0, -h]])
    stacked = (points_mm[None, :, :] + offsets[:, None, :]).reshape(-1, 3)
    vals = eval_batch(node, stacked, device=device).reshape(6, n)
    gx = (vals[0] - vals[1]) / (2 * h)
    gy = (vals[2] - vals[3]) / (2 * h)
    gz = (vals[4] - vals[5]) / (2 * h)
    return np.stack([gx, gy, gz], axis=1)


def klassificera_och_evaluera_ikarus(meta, node, device="cuda:0"):
    """Block-sparse field of a tree on the grid `meta` (a field.GridMeta):
    classification at the eight block corners (FE/faltkarna_v1.py:128-164)
    on the device, ascending compaction of the surface blocks on the host,
    tiles of the active blocks on the device (FE/faltkarna_v1.py:166-190).
    Returns a field.SparseField."""
    import torch
    from .field import SparseField, _stats
    prog = _program(node)
    dev = torch.device(device)
    ins, par, grids, gdata = prog.device(dev)
    L = _lib.lib()
    P = _lib.ptr
    st = _lib.stream_handle(None)
    opm = np.array([meta.x0, meta.y0, meta.z0, meta.pitch, meta.margin_mm], np.float32)
    blocks = np.array([meta.n_bx, meta.n_by, meta.n_bz], np.int32)
    kind_d = torch.empty(meta.n_blocks, dtype=torch.int32, device=dev)
    cma_d = torch.empty(meta.n_blocks, dtype=torch.float32, device=dev)
    _check(L.fe_ikarus_classify(P(ins), len(prog.ins), P(par), P(grids), P(gdata), opm.ctypes.data, blocks.ctypes.data,
                                meta.block, P(kind_d), P(cma_d), st), "ikarus classify")
    kind = kind_d.cpu().numpy()
    cma = cma_d.cpu().numpy()
    active_ids = np.where(kind == 0)[0].astype(np.int32)
    B = meta.block
    if len(active_ids):
        ids_d = torch.from_numpy(active_ids).to(dev)
        tiles_d = torch.empty(len(active_ids) * B ** 3, dtype=torch.float32, device=dev)
        _check(L.fe_ikarus_tiles(P(ins), len(prog.ins), P(par), P(grids), P(gdata), opm.ctypes.data,
                                 blocks.ctypes.data, B, P(ids_d), len(active_ids), P(tiles_d), st), "ikarus tiles")
        tiles = tiles_d.cpu().numpy().reshape(-1, B, B, B)
    else:
        tiles = np.zeros((0, B, B, B), np.float32)
    return SparseField(meta=meta, kind=kind, active_ids=active_ids, tiles=tiles, corner_min_abs=cma,
                       stats=_stats(meta, kind, len(active_ids)))


# ------------------------------------------------------------------ grid-leaf pairs and band trees
_LEAF_TABLES = {}
_LEAF_TABLES_MAX = 32


def _leaf_key(node):
    """Identity of a grid leaf's data: (path, key, mode, iso, pitch override)
    for an .npz leaf (the reference's upload-cache key,
    FE/ikarus_v1/eval_warp.py:441, plus the pitch override), the array object
    for an in-memory leaf."""
    if node.get("op") != "gridfield":
        raise ValueError("a grid leaf (op 'gridfield') is required")
    if "values" in node:
        return ("mem", id(node["values"]), tuple(node["origin"]), node["pitch"], node["mode"], node["iso"])
    return ("npz", node["npz_path"], node["key"], node["mode"], node["iso"], node.get("pitch_override"))


def leaf_table(leaves, device="cuda:0"):
    """Device grid records (uint8 view of 48-byte records) and fp32 samples of
    a list of grid leaves, uploaded once per (leaves, device) and reused
    (FE/ikarus_v1/eval_warp.py:433-469)."""
    import torch
    dev = torch.device(device)
    key = (tuple(_leaf_key(n) for n in leaves), str(dev))
    hit = _LEAF_TABLES.get(key)
    if hit is not None:
        return hit[0], hit[1]
    if not leaves:
        raise ValueError("at least one grid leaf is required")
    recs, chunks, off = [], [], 0
    for n in leaves:
        rec, arr = _grid_record(n, off)
        recs.append(rec)
        chunks.append(arr.reshape(-1))
        off += arr.size
    grids = torch.from_numpy(np.array(recs, GRID_REC).view(np.uint8)).to(dev)
    gdata = torch.from_numpy(np.concatenate(chunks).astype(np.float32)).to(dev)
    if len(_LEAF_TABLES) >= _LEAF_TABLES_MAX:
        _LEAF_TABLES.pop(next(iter(_LEAF_TABLES)))
    # the leaves are kept with the entry so that in-memory arrays keep their ids
    _LEAF_TABLES[key] = (grids, gdata, list(leaves))
    return grids, gdata


def eval_batch_multi_pair(points_mm, pair_of_point, leaf_a_of_pair, leaf_b_of_pair, grid_leaves_all,
                          device="cuda:0"):
    """Hard intersection of two grid leaves per point, the pair chosen per
    point, in one launch (FE/ikarus_v1/eval_warp.py:545-560): (N,) fp32
    max(leaf a, leaf b) at the points cast to fp32."""
    import torch
    pts = np.ascontiguousarray(np.asarray(points_mm, np.float64).astype(np.float32)).reshape(-1, 3)
    pop = np.ascontiguousarray(pair_of_point, np.int32).reshape(-1)
    la = np.ascontiguousarray(leaf_a_of_pair, np.int32).reshape(-1)
    lb = np.ascontiguousarray(leaf_b_of_pair, np.int32).reshape(-1)
    nl = len(grid_leaves_all)
    if pop.shape[0] != pts.shape[0]:
        raise ValueError("pair_of_point needs one entry per point")
    if len(la) != len(lb) or len(la) < 1:
        raise ValueError("leaf_a_of_pair and leaf_b_of_pair need the same, nonzero length")
    if pop.size and (pop.min() < 0 or pop.max() >= len(la)):
        raise ValueError("pair index out of range")
    if la.min() < 0 or lb.min() < 0 or max(la.max(), lb.max()) >= nl:
        raise ValueError("leaf index out of range")
    dev = torch.device(device)
    grids, gdata = leaf_table(grid_leaves_all, dev)
    n = pts.shape[0]
    out = torch.empty(n, dtype=torch.float32, device=dev)
    if n:
        t_pts = torch.from_numpy(pts).to(dev)
        t_pop = torch.from_numpy(pop).to(dev)
        t_la = torch.from_numpy(la).to(dev)
        t_lb = torch.from_numpy(lb).to(dev)
        P = _lib.ptr
        _check(_lib.lib().fe_ikarus_pairs_eval(P(grids), nl, P(gdata), P(t_pts), P(t_pop), P(t_la), P(t_lb), len(la),
                                               n, P(out), _lib.stream_handle(None)), "ikarus pairs eval")
    return out.cpu().numpy()


def band_tree_expr(leaf_node, axis, bands):
    """The tree of a band plan (FE/ikarus_v1/eval_warp.py:700-715):
    union over bands of intersect(axis_project(leaf, axis, zs),
    intersect(halfspace(+e, z1), halfspace(-e, -z0)))."""
    def slab(z0, z1):
        n = [0.0, 0.0, 0.0]
        n[axis] = 1.0
        nm = [0.0, 0.0, 0.0]
        nm[axis] = -1.0
        return intersect(halfspace(n, z1), halfspace(nm, -z0))
    node = None
    for b in bands:
        piece = intersect(axis_project(leaf_node, axis, b["zs"]), slab(b["z0"], b["z1"]))
        node = piece if node is None else union(node, piece)
    return node


class BandTreeEvaluator:
    """A grid leaf and the band-tree pass, the plan swappable per call
    (FE/ikarus_v1/eval_warp.py:659-697).  The leaf is uploaded once; a plan is
    three fp32 arrays (zs, z0, z1), cached per plan.  Values equal the
    interpreter's on band_tree_expr(leaf, axis, bands) bit for bit."""

    def __init__(self, leaf_node, axis, device="cuda:0"):
        import torch
        if int(axis) not in (0, 1, 2):
            raise ValueError("axis must be 0, 1 or 2")
        self.leaf = leaf_node
        self.axis = int(axis)
        self.device = torch.device(device)
        self.grid_args = leaf_table([leaf_node], self.device)
        self._plan_cache = {}

    def _plan_arrays(self, bands):
        import torch
        key = tuple((float(b["zs"]), float(b["z0"]), float(b["z1"])) for b in bands)
        got = self._plan_cache.get(key)
        if got is None:
            plan = np.array([[b["zs"] for b in bands], [b["z0"] for b in bands], [b["z1"] for b in bands]],
                            np.float32).reshape(3, len(bands))
            got = (torch.from_numpy(np.ascontiguousarray(plan)).to(self.device), len(bands))
            self._plan_cache[key] = got
        return got

    def eval(self, points_mm, bands):
        """(N,) fp32 values of the band tree at the points (cast to fp32)."""
        import torch
        plan, nb = self._plan_arrays(bands)
        pts = np.ascontiguousarray(np.asarray(points_mm, np.float64).astype(np.float32)).reshape(-1, 3)
        n = pts.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=self.device)
        if n:
            t_pts = torch.from_numpy(pts).to(self.device)
            grids, gdata = self.grid_args
            P = _lib.ptr
            base = plan.data_ptr()
            _check(_lib.lib().fe_ikarus_band_eval(P(grids), P(gdata), self.axis, base, base + 4 * nb, base + 8 * nb,
                                                  nb, P(t_pts), n, P(out), _lib.stream_handle(None)),
                   "ikarus band eval")
        return out.cpu().numpy()
