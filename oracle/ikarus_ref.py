"""CPU oracle for IKARUS expression trees.  TEST INFRASTRUCTURE ONLY: the
tests and the benchmark's CPU leg use it as the checker; the product path
never does.

Two restatements, both walking the tree itself (not the package's compiled
program):

* `eval_f32(node, pts)`: the fp32 arithmetic of the per-node functions the
  reference generates (FE/ikarus_v1/eval_warp.py:175-320 with the grid sampler
  of :92-148; rect_frustum as FE/ikarus_v1/eval_warp_uniform_v2.py:180-214):
  parameters rounded to fp32 from the node's doubles, every operation one
  fp32 rounding in the generated order, Warp's min/max (a<b?a:b, a>b?a:b),
  clamp (min(max(lo, x), hi)) and lerp (a*(1-t) + b*t);
* `eval_f64(node, pts)`: the pure-Python ground truth eval_sdf_py
  (FE/ikarus_v1/expr.py:263-385) in float64, vectorised over points.

`classify_f32` / `tiles_f32` restate the block classification and tile
evaluation of FE/faltkarna_v1.py:128-190 on top of eval_f32.
"""
import math

import numpy as np

F = np.float32
AX = {"x": 0, "y": 1, "z": 2}


def inverse_affine(m):
    """FE/ikarus_v1/expr.py:57-83 (cofactors, then the translation)."""
    R = [[m[i][j] for j in range(3)] for i in range(3)]
    t = [m[i][3] for i in range(3)]
    det = (R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1])
           - R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0])
           + R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]))
    if abs(det) < 1e-15:
        raise ValueError("singular transform")
    inv = [[0.0] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            r1, r2, c1, c2 = (i + 1) % 3, (i + 2) % 3, (j + 1) % 3, (j + 2) % 3
            inv[j][i] = (R[r1][c1] * R[r2][c2] - R[r1][c2] * R[r2][c1]) / det
    out = [row[:] + [0.0] for row in inv] + [[0.0, 0.0, 0.0, 1.0]]
    out[0][3] = -(inv[0][0] * t[0] + inv[0][1] * t[1] + inv[0][2] * t[2])
    out[1][3] = -(inv[1][0] * t[0] + inv[1][1] * t[1] + inv[1][2] * t[2])
    out[2][3] = -(inv[2][0] * t[0] + inv[2][1] * t[1] + inv[2][2] * t[2])
    return out


# ------------------------------------------------------------------ fp32
def _wmin(a, b):
    return np.where(a < b, a, b).astype(F)


def _wmax(a, b):
    return np.where(a > b, a, b).astype(F)


def _lerp(a, b, t):
    return (a * (F(1.0) - t) + b * t).astype(F)


def _grid_leaf(node):
    if "values" in node:
        return (np.asarray(node["values"], F), node["origin"], node["pitch"], node["iso"], node["mode"] == "sdf")
    with np.load(node["npz_path"], allow_pickle=False) as d:
        arr = np.asarray(d[node["key"]], F)
        origin = d["global_origin_mm"] if "global_origin_mm" in d else np.zeros(3)
        pitch = node["pitch_override"] or float(d["pitch_mm"]) if "pitch_mm" in d else 1.0
    return arr, [float(x) for x in origin], pitch, node["iso"], node["mode"] == "sdf"


def _sample(node, px, py, pz):
    arr, origin, pitch, iso, is_sdf = _grid_leaf(node)
    nx, ny, nz = arr.shape
    h = F(pitch)
    fx = (px - F(origin[0])) / h
    fy = (py - F(origin[1])) / h
    fz = (pz - F(origin[2])) / h
    out = np.full(px.shape, F(1.0e6))
    inside = ~((fx < 0) | (fy < 0) | (fz < 0) | (fx > F(nx - 1)) | (fy > F(ny - 1)) | (fz > F(nz - 1)))
    fx, fy, fz = fx[inside], fy[inside], fz[inside]
    ix, iy, iz = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64), np.floor(fz).astype(np.int64)
    ix1, iy1, iz1 = np.minimum(ix + 1, nx - 1), np.minimum(iy + 1, ny - 1), np.minimum(iz + 1, nz - 1)
    tx, ty, tz = (fx - ix.astype(F)).astype(F), (fy - iy.astype(F)).astype(F), (fz - iz.astype(F)).astype(F)
    c00 = _lerp(arr[ix, iy, iz], arr[ix1, iy, iz], tx)
    c10 = _lerp(arr[ix, iy1, iz], arr[ix1, iy1, iz], tx)
    c01 = _lerp(arr[ix, iy, iz1], arr[ix1, iy, iz1], tx)
    c11 = _lerp(arr[ix, iy1, iz1], arr[ix1, iy1, iz1], tx)
    val = _lerp(_lerp(c00, c10, ty), _lerp(c01, c11, ty), tz)
    out[inside] = val if is_sdf else ((F(iso) - val) * h).astype(F)
    return out


def _axial(axis, lx, ly, lz):
    if axis == "z":
        return np.sqrt(lx * lx + ly * ly), lz
    if axis == "y":
        return np.sqrt(lx * lx + lz * lz), ly
    return np.sqrt(ly * ly + lz * lz), lx


def _round2(qx, qy):
    ox, oy = _wmax(qx, F(0)), _wmax(qy, F(0))
    return (np.sqrt(ox * ox + oy * oy) + _wmin(_wmax(qx, qy), F(0))).astype(F)


def _f32(node, px, py, pz):
    op = node["op"]
    if op == "halfspace":
        nx, ny, nz = (F(v) for v in node["normal"])
        return px * nx + py * ny + pz * nz - F(node["offset"])
    if op == "sphere":
        cx, cy, cz = (F(v) for v in node["center"])
        dx, dy, dz = px - cx, py - cy, pz - cz
        return np.sqrt(dx * dx + dy * dy + dz * dz) - F(node["radius"])
    if op == "box":
        cx, cy, cz = (F(v) for v in node["center"])
        hx, hy, hz = (F(v) for v in node["half_extents"])
        qx, qy, qz = np.abs(px - cx) - hx, np.abs(py - cy) - hy, np.abs(pz - cz) - hz
        ox, oy, oz = _wmax(qx, F(0)), _wmax(qy, F(0)), _wmax(qz, F(0))
        return np.sqrt(ox * ox + oy * oy + oz * oz) + _wmin(_wmax(qx, _wmax(qy, qz)), F(0))
    if op in ("cylinder", "cone"):
        cx, cy, cz = (F(v) for v in node["center"])
        rad, ax = _axial(node["axis"], px - cx, py - cy, pz - cz)
        h2 = F(node["height"] / 2.0)
        if op == "cylinder":
            return _round2(rad - F(node["radius"]), np.abs(ax) - h2)
        t = (ax + h2) / F(node["height"])
        t = _wmin(_wmax(F(0), t), F(1))
        r1, r2 = F(node["radius1"]), F(node["radius2"])
        r_at = r1 + (r2 - r1) * t
        return _round2(rad - r_at, np.abs(ax) - h2)
    if op == "rect_frustum":
        z0, z1 = node["z0"], node["z1"]
        kx = (node["hx1"] - node["hx0"]) / (z1 - z0)
        ky = (node["hy1"] - node["hy0"]) / (z1 - z0)
        kcx = (node["cx1"] - node["cx0"]) / (z1 - z0)
        kcy = (node["cy1"] - node["cy0"]) / (z1 - z0)
        nzhi, chi = F(-(kx + kcx)), F(node["hx0"] + node["cx0"] - (kx + kcx) * z0)
        nzlo, clo = F(-(kx - kcx)), F(node["hx0"] - node["cx0"] - (kx - kcx) * z0)
        ynzhi, ychi = F(-(ky + kcy)), F(node["hy0"] + node["cy0"] - (ky + kcy) * z0)
        ynzlo, yclo = F(-(ky - kcy)), F(node["hy0"] - node["cy0"] - (ky - kcy) * z0)
        one, mone = F(1.0), F(-1.0)
        d_zlo, d_zhi = F(z0) - pz, pz - F(z1)
        d_xhi = (px + nzhi * pz - chi) / np.sqrt(one * one + nzhi * nzhi)
        d_xlo = (-px + nzlo * pz - clo) / np.sqrt(mone * mone + nzlo * nzlo)
        d_yhi = (py + ynzhi * pz - ychi) / np.sqrt(one * one + ynzhi * ynzhi)
        d_ylo = (-py + ynzlo * pz - yclo) / np.sqrt(mone * mone + ynzlo * ynzlo)
        return _wmax(_wmax(d_zlo, d_zhi), _wmax(_wmax(d_xhi, d_xlo), _wmax(d_yhi, d_ylo)))
    if op == "gridfield":
        return _sample(node, px, py, pz)
    if op == "transform":
        m = [[F(v) for v in row] for row in inverse_affine(node["mat4"])]
        lx = m[0][0] * px + m[0][1] * py + m[0][2] * pz + m[0][3]
        ly = m[1][0] * px + m[1][1] * py + m[1][2] * pz + m[1][3]
        lz = m[2][0] * px + m[2][1] * py + m[2][2] * pz + m[2][3]
        return _f32(node["child"], lx, ly, lz)
    if op == "axis_project":
        p = [px, py, pz]
        p[node["axis"]] = np.full(px.shape, F(node["ref"]))
        return _f32(node["child"], *p)
    if op == "offset":
        return _f32(node["child"], px, py, pz) - F(node["delta"])
    if op == "scale":
        s = F(node["s"])
        return _f32(node["child"], px / s, py / s, pz / s) * s
    if op in ("union", "intersect", "subtract"):
        da = _f32(node["a"], px, py, pz).astype(F)
        db = _f32(node["b"], px, py, pz).astype(F)
        k = F(node["k"])
        if op == "union":
            hard = _wmin(da, db)
            if k <= 0:
                return hard
            h = _wmax(F(0), k - np.abs(da - db)) / k
            return hard - h * h * k * F(0.25)
        if op == "intersect":
            hard = _wmax(da, db)
            if k <= 0:
                return hard
            h = _wmax(F(0), k - np.abs(da - db)) / k
            return hard + h * h * k * F(0.25)
        hard = _wmax(da, -db)
        if k <= 0:
            return hard
        h = _wmax(F(0), k - np.abs(da + db)) / k
        return hard + h * h * k * F(0.25)
    raise ValueError(f"unknown op {op!r}")


def eval_f32(node, pts):
    p = np.asarray(pts, dtype=np.float64).astype(F).reshape(-1, 3)
    with np.errstate(all="ignore"):
        out = _f32(node, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy())
    return np.broadcast_to(np.asarray(out, F), (len(p),)).copy()


# ------------------------------------------------------------------ float64 ground truth
def _f64(node, x, y, z):
    op = node["op"]
    if op == "halfspace":
        nx, ny, nz = node["normal"]
        return x * nx + y * ny + z * nz - node["offset"]
    if op == "sphere":
        cx, cy, cz = node["center"]
        return np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) - node["radius"]
    if op == "box":
        cx, cy, cz = node["center"]
        hx, hy, hz = node["half_extents"]
        qx, qy, qz = np.abs(x - cx) - hx, np.abs(y - cy) - hy, np.abs(z - cz) - hz
        outside = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
        return outside + np.minimum(np.maximum(qx, np.maximum(qy, qz)), 0)
    if op in ("cylinder", "cone"):
        cx, cy, cz = node["center"]
        lx, ly, lz = x - cx, y - cy, z - cz
        axis = node["axis"]
        if axis == "z":
            rad, ax = np.sqrt(lx * lx + ly * ly), lz
        elif axis == "y":
            rad, ax = np.sqrt(lx * lx + lz * lz), ly
        else:
            rad, ax = np.sqrt(ly * ly + lz * lz), lx
        h = node["height"]
        if op == "cylinder":
            qx, qy = rad - node["radius"], np.abs(ax) - h / 2.0
        else:
            t = np.clip((ax + h / 2.0) / h, 0.0, 1.0)
            qx = rad - (node["radius1"] + (node["radius2"] - node["radius1"]) * t)
            qy = np.abs(ax) - h / 2.0
        return np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2) + np.minimum(np.maximum(qx, qy), 0)
    if op == "rect_frustum":
        z0, z1 = node["z0"], node["z1"]
        kx = (node["hx1"] - node["hx0"]) / (z1 - z0)
        ky = (node["hy1"] - node["hy0"]) / (z1 - z0)
        kcx = (node["cx1"] - node["cx0"]) / (z1 - z0)
        kcy = (node["cy1"] - node["cy0"]) / (z1 - z0)

        def plane(pv, nx, nz, c):
            return (nx * pv + nz * z - c) / math.sqrt(nx * nx + nz * nz)
        faces = [z0 - z, z - z1,
                 plane(x, 1.0, -(kx + kcx), node["hx0"] + node["cx0"] - (kx + kcx) * z0),
                 plane(x, -1.0, -(kx - kcx), node["hx0"] - node["cx0"] - (kx - kcx) * z0),
                 plane(y, 1.0, -(ky + kcy), node["hy0"] + node["cy0"] - (ky + kcy) * z0),
                 plane(y, -1.0, -(ky - kcy), node["hy0"] - node["cy0"] - (ky - kcy) * z0)]
        return np.maximum.reduce(faces)
    if op == "gridfield":
        raise NotImplementedError("the reference has no float64 grid sampler")
    if op == "transform":
        m = inverse_affine(node["mat4"])
        return _f64(node["child"], m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3],
                    m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3],
                    m[2][0] * x + m[2][1] * y + m[2][2] * z + m[2][3])
    if op == "axis_project":
        p = [x, y, z]
        p[node["axis"]] = np.full(x.shape, node["ref"])
        return _f64(node["child"], *p)
    if op == "offset":
        return _f64(node["child"], x, y, z) - node["delta"]
    if op == "scale":
        s = node["s"]
        return _f64(node["child"], x / s, y / s, z / s) * s
    if op in ("union", "intersect", "subtract"):
        da, db = _f64(node["a"], x, y, z), _f64(node["b"], x, y, z)
        k = node["k"]
        if op == "union":
            if k <= 0:
                return np.minimum(da, db)
            h = np.maximum(k - np.abs(da - db), 0.0) / k
            return np.minimum(da, db) - h * h * k * 0.25
        if op == "intersect":
            if k <= 0:
                return np.maximum(da, db)
            h = np.maximum(k - np.abs(da - db), 0.0) / k
            return np.maximum(da, db) + h * h * k * 0.25
        if k <= 0:
            return np.maximum(da, -db)
        h = np.maximum(k - np.abs(da + db), 0.0) / k
        return np.maximum(da, -db) + h * h * k * 0.25
    raise ValueError(f"unknown op {op!r}")


def eval_f64(node, pts):
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    out = _f64(node, p[:, 0], p[:, 1], p[:, 2])
    return np.broadcast_to(np.asarray(out, np.float64), (len(p),)).copy()


# ------------------------------------------------------------------ block-sparse build
def _block_index(meta):
    nb = meta["nbx"] * meta["nby"] * meta["nbz"]
    b = np.arange(nb)
    return b // (meta["nby"] * meta["nbz"]), (b // meta["nbz"]) % meta["nby"], b % meta["nbz"]


def classify_f32(node, meta):
    """(kind int32, corner_min_abs fp32) per block (FE/faltkarna_v1.py:128-164):
    corners at x0 + f32(bi*B)*h + {0, f32(B)*h}; mixed signs or min|v| <=
    margin -> 0, all negative -> -1, else +1."""
    B = meta["B"]
    h = F(meta["pitch"])
    side = F(B) * h
    bi, bj, bk = _block_index(meta)
    ox = F(meta["x0"]) + (bi * B).astype(F) * h
    oy = F(meta["y0"]) + (bj * B).astype(F) * h
    oz = F(meta["z0"]) + (bk * B).astype(F) * h
    vals = []
    for c in range(8):
        x = ox + F(c & 1) * side
        y = oy + F((c >> 1) & 1) * side
        z = oz + F((c >> 2) & 1) * side
        vals.append(eval_f32(node, np.stack([x, y, z], 1)))
    v = np.stack(vals, 1)
    pos = (v >= 0).any(1)
    neg = (v < 0).any(1)
    mn = np.abs(v).min(1).astype(F)
    kind = np.where(pos & neg, 0, np.where(mn <= F(meta["margin"]), 0, np.where(neg, -1, 1))).astype(np.int32)
    return kind, mn


def tiles_f32(node, meta, ids):
    B = meta["B"]
    h = F(meta["pitch"])
    ids = np.asarray(ids, np.int64)
    bi, bj, bk = ids // (meta["nby"] * meta["nbz"]), (ids // meta["nbz"]) % meta["nby"], ids % meta["nbz"]
    li, lj, lk = np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")
    x = F(meta["x0"]) + (bi[:, None, None, None] * B + li[None]).astype(F) * h
    y = F(meta["y0"]) + (bj[:, None, None, None] * B + lj[None]).astype(F) * h
    z = F(meta["z0"]) + (bk[:, None, None, None] * B + lk[None]).astype(F) * h
    return eval_f32(node, np.stack([x.ravel(), y.ravel(), z.ravel()], 1)).reshape(len(ids), B, B, B)


# ------------------------------------------------------------------ hierarchical volume (FE/ikarus_v1/queries.py)
def grid_points(lo, hi, pitch):
    """FE/ikarus_v1/queries.py:120-131: samples lo + pitch*i, C order (float64)."""
    n = [max(int(math.ceil((hi[a] - lo[a]) / pitch)) + 1, 1) for a in range(3)]
    axes = [lo[a] + pitch * np.arange(n[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1), tuple(n)


def grid_volume(node, lo, hi, coarse_pitch, refine_factor):
    """The coarse and refine stages of FE/ikarus_v1/queries.py:211-264 at given
    pitches, on eval_f32 (points cast to fp32 as eval_batch does)."""
    pts_c, shape = grid_points(lo, hi, coarse_pitch)
    vals_c = eval_f32(node, pts_c)
    diag_half = 0.5 * coarse_pitch * math.sqrt(3.0)
    inside = vals_c < F(-diag_half)
    boundary = np.abs(vals_c) <= F(diag_half)
    out = dict(n_fully_inside=int(inside.sum()), n_boundary=int(boundary.sum()), grid_shape=shape,
               volume_mm3=float(inside.sum()) * coarse_pitch ** 3)
    fine_in = fine_bd = 0
    if out["n_boundary"] and refine_factor > 1:
        half = coarse_pitch / 2.0
        offs = (np.arange(refine_factor) + 0.5) / refine_factor * coarse_pitch - half
        OX, OY, OZ = np.meshgrid(offs, offs, offs, indexing="ij")
        local = np.stack([OX.ravel(), OY.ravel(), OZ.ravel()], axis=1)
        fine = (pts_c[boundary][:, None, :] + local[None, :, :]).reshape(-1, 3)
        fv = eval_f32(node, fine)
        fine_in = int((fv < F(0)).sum())
        fine_bd = int((np.abs(fv) <= F(0.5 * (coarse_pitch / refine_factor) * math.sqrt(3.0))).sum())
        out["volume_mm3"] += float(fine_in) * (coarse_pitch / refine_factor) ** 3
    out.update(n_fine_inside=fine_in, n_fine_boundary=fine_bd)
    return out


# ------------------------------------------------------------------ grid-leaf pairs, band trees, pair sweeps
def pairs_f32(points, pair_of_point, leaf_a, leaf_b, leaves):
    """FE/ikarus_v1/eval_warp.py:504-560: per point, max(leaf a, leaf b) of its
    pair with the selection max, at the point cast to fp32."""
    p = np.asarray(points, np.float64).astype(F).reshape(-1, 3)
    pop = np.asarray(pair_of_point)
    out = np.empty(len(p), F)
    for k in np.unique(pop):
        sel = pop == k
        x, y, z = p[sel, 0], p[sel, 1], p[sel, 2]
        out[sel] = _wmax(_sample(leaves[leaf_a[k]], x, y, z), _sample(leaves[leaf_b[k]], x, y, z))
    return out


def band_f32(leaf, axis, bands, points):
    """FE/ikarus_v1/eval_warp.py:601-657: min over bands of max(leaf at the
    point with p[axis] := zs, max(p[axis] - z1, z0 - p[axis])), folded from
    1e30 (fp32 plan values)."""
    p = np.asarray(points, np.float64).astype(F).reshape(-1, 3)
    pax = p[:, axis]
    best = np.full(len(p), F(1.0e30))
    for b in bands:
        q = p.copy()
        q[:, axis] = F(b["zs"])
        body = _sample(leaf, q[:, 0], q[:, 1], q[:, 2])
        slab = _wmax((pax - F(b["z1"])).astype(F), (F(b["z0"]) - pax).astype(F))
        best = _wmin(best, _wmax(body, slab))
    return best
