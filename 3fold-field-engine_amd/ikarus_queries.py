"""Hierarchical volume queries over IKARUS trees (FE/ikarus_v1/queries.py:138-277).

Every pass is one device kernel over an implicit grid (csrc/ikarus.hip
`k_ik_grid`, `k_ik_refine`): samples are generated on the device from
(lo, pitch, shape) with the f64 arithmetic of a NumPy grid and rounded to fp32
as eval_batch would, classified in registers and counted by block-level
reductions; only the undecided cells' indices are kept for the refinement, so
no point array crosses the bus.

`grid_volume` runs the two passes at a given coarse pitch and refinement
factor.  `intrusion_volume(a, b)` sizes them from a point budget: the coarse
grid takes half of it over the overlap of the operands' boxes, and the
refinement factor is the largest (up to `refine_factor`) whose samples fit
the rest.  The reference sizes its passes from a wall-clock budget through a
calibrated cost model; that scheduling is not reproduced here (a point budget
gives the same passes on every run).
"""
import math

import numpy as np

from . import _lib
from . import ikarus as E


def _leaf_box(n):
    """Local (lo, hi) of a leaf, or None for a halfspace (unbounded)."""
    op = n["op"]
    if op == "sphere":
        c, r = np.asarray(n["center"], np.float64), float(n["radius"])
        return c - r, c + r
    if op == "box":
        c, h = np.asarray(n["center"], np.float64), np.abs(np.asarray(n["half_extents"], np.float64))
        return c - h, c + h
    if op in ("cylinder", "cone"):
        c = np.asarray(n["center"], np.float64)
        r = max(abs(float(n.get(k, 0.0))) for k in ("radius", "radius1", "radius2"))
        ext = np.full(3, r)
        ext[E.AXES[n["axis"]]] = abs(float(n["height"])) / 2.0
        return c - ext, c + ext
    if op == "rect_frustum":
        hx, hy = max(n["hx0"], n["hx1"]), max(n["hy0"], n["hy1"])
        lo = np.array([min(n["cx0"], n["cx1"]) - hx, min(n["cy0"], n["cy1"]) - hy, min(n["z0"], n["z1"])])
        hi = np.array([max(n["cx0"], n["cx1"]) + hx, max(n["cy0"], n["cy1"]) + hy, max(n["z0"], n["z1"])])
        return lo, hi
    if op == "gridfield":
        arr, origin, pitch, _, _ = E._grid_samples(n)
        lo = np.asarray(origin, np.float64)
        return lo, lo + pitch * (np.asarray(arr.shape) - 1)
    if op == "halfspace":
        return None
    raise ValueError(f"unknown node {op!r}")


def _map_box(lo, hi, xf):
    """World box of a local box under the affine map xf (its 8 corners)."""
    corners = np.array([[(lo, hi)[i >> a & 1][a] for a in range(3)] + [1.0] for i in range(8)])
    world = corners @ xf.T
    return world[:, :3].min(0), world[:, :3].max(0)


def bbox_of(node, pad=1.0):
    """Axis-aligned world box of a tree, padded by `pad` on every side: the
    union of its leaves' boxes (a cylinder or cone along its own axis),
    mapped through the enclosing transforms; offset, scale and axis_project
    keep their child's box and halfspaces add nothing."""
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    stack = [(node, np.eye(4))]
    while stack:
        n, xf = stack.pop()
        op = n["op"]
        if op == "transform":
            stack.append((n["child"], xf @ np.asarray(n["mat4"], np.float64)))
        elif op in ("offset", "scale", "axis_project"):
            stack.append((n["child"], xf))
        elif op in ("union", "intersect", "subtract"):
            stack.extend(((n["a"], xf), (n["b"], xf)))
        else:
            box = _leaf_box(n)
            if box is not None:
                blo, bhi = _map_box(*box, xf)
                lo, hi = np.minimum(lo, blo), np.maximum(hi, bhi)
    if not np.all(np.isfinite(lo)):
        raise ValueError("tree has no bounded leaf")
    return lo - pad, hi + pad


def grid_shape(lo, hi, pitch):
    """Samples per axis of the grid lo + pitch*i reaching hi."""
    return tuple(max(int(math.ceil((hi[a] - lo[a]) / pitch)) + 1, 1) for a in range(3))


class _Passes:
    """Device state of the passes of one query: the program and scratch."""

    def __init__(self, node, device):
        import torch
        self.torch = torch
        self.prog = node if isinstance(node, E.Program) else E.compile_tree(node)
        self.dev = torch.device(device)
        self.ins, self.par, self.grids, self.gdata = self.prog.device(self.dev)
        self.counts = torch.zeros(2, dtype=torch.int64, device=self.dev)
        self.cursor = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.cells = None

    def grid(self, lo, pitch, shape, below, band, keep_cells):
        t = self.torch
        L = _lib.lib()
        P = _lib.ptr
        n = int(np.prod(shape))
        self.counts.zero_()
        cells = cursor = None
        if keep_cells:
            if self.cells is None or self.cells.numel() < n:
                self.cells = t.empty(n, dtype=t.int32, device=self.dev)
            self.cursor.zero_()
            cells, cursor = P(self.cells), P(self.cursor)
        lo_a = np.ascontiguousarray(lo, np.float64)
        sh = np.ascontiguousarray(shape, np.int32)
        st = L.fe_ikarus_grid(P(self.ins), len(self.prog.ins), P(self.par), P(self.grids), P(self.gdata),
                              lo_a.ctypes.data, float(pitch), sh.ctypes.data, float(np.float32(below)),
                              float(np.float32(band)), P(self.counts), cells, cursor, _lib.stream_handle(None))
        if st:
            raise RuntimeError(f"grid pass failed with status {st}")
        c = self.counts.cpu().numpy()
        return int(c[0]), int(c[1])

    def refine(self, lo, pitch, shape, ncells, r, band):
        L = _lib.lib()
        P = _lib.ptr
        self.counts.zero_()
        lo_a = np.ascontiguousarray(lo, np.float64)
        sh = np.ascontiguousarray(shape, np.int32)
        st = L.fe_ikarus_refine(P(self.ins), len(self.prog.ins), P(self.par), P(self.grids), P(self.gdata),
                                lo_a.ctypes.data, float(pitch), sh.ctypes.data, P(self.cells), int(ncells), int(r),
                                float(np.float32(band)), P(self.counts), _lib.stream_handle(None))
        if st:
            raise RuntimeError(f"refine pass failed with status {st}")
        c = self.counts.cpu().numpy()
        return int(c[0]), int(c[1])


def grid_volume(node, lo, hi, coarse_pitch, refine_factor, device="cuda:0", passes=None):
    """Volume of {v < 0} over the box [lo, hi] at a given coarse pitch and
    refinement factor (the coarse and refine stages of
    FE/ikarus_v1/queries.py:211-264): coarse cells with v below minus the
    cell's half diagonal count whole, cells within it are refined into r^3
    samples each, and the fine samples still within the fine half diagonal
    make up the declared uncertainty.  Returns a dict with the counts."""
    p = passes or _Passes(node, device)
    shape = grid_shape(lo, hi, coarse_pitch)
    diag_half = 0.5 * coarse_pitch * math.sqrt(3.0)
    n_inside, n_boundary = p.grid(lo, coarse_pitch, shape, -diag_half, diag_half, keep_cells=refine_factor > 1)
    vol = float(n_inside) * coarse_pitch ** 3
    fine_in = fine_bd = 0
    if n_boundary > 0 and refine_factor > 1:
        fine_diag_half = 0.5 * (coarse_pitch / refine_factor) * math.sqrt(3.0)
        fine_in, fine_bd = p.refine(lo, coarse_pitch, shape, n_boundary, refine_factor, fine_diag_half)
        vol += float(fine_in) * (coarse_pitch / refine_factor) ** 3
        unc = fine_bd * (coarse_pitch / refine_factor) ** 3
    else:
        unc = n_boundary * coarse_pitch ** 3
    return dict(volume_mm3=vol, uncertainty_mm3=unc, grid_shape=shape, n_fully_inside=n_inside,
                n_boundary=n_boundary, n_fine_inside=fine_in, n_fine_boundary=fine_bd,
                n_points=int(np.prod(shape)) + n_boundary * refine_factor ** 3 * (refine_factor > 1))


def intrusion_volume(exprA, exprB, max_points=4_000_000, refine_factor=4, min_pitch_mm=0.25, device="cuda:0"):
    """Volume of A intersect B (FE/ikarus_v1/queries.py:138-277) within a
    budget of `max_points` samples: the coarse grid gets half of the budget
    over the overlap of the operands' boxes (pitch at least min_pitch_mm), the
    refinement factor is the largest r <= refine_factor whose r^3 samples per
    undecided cell fit what is left (r = 1: no refinement, the undecided cells
    are the uncertainty).  Returns grid_volume's dict plus the pitches."""
    alo, ahi = bbox_of(exprA)
    blo, bhi = bbox_of(exprB)
    lo, hi = np.maximum(alo, blo), np.minimum(ahi, bhi)
    if np.any(hi <= lo):
        return dict(volume_mm3=0.0, uncertainty_mm3=0.0, n_points=0, coarse_pitch_mm=None, refine_factor=0,
                    empty_overlap=True)
    passes = _Passes(E.intersect(exprA, exprB), device)
    pitch = max((float(np.prod(hi - lo)) / max(max_points // 2, 1)) ** (1.0 / 3.0), float(min_pitch_mm))
    shape = grid_shape(lo, hi, pitch)
    half = 0.5 * pitch * math.sqrt(3.0)
    n_in, n_bd = passes.grid(lo, pitch, shape, -half, half, keep_cells=True)
    left = max_points - int(np.prod(shape))
    r = max(refine_factor, 1)
    while r > 1 and n_bd * r ** 3 > left:
        r -= 1
    out = dict(volume_mm3=n_in * pitch ** 3, uncertainty_mm3=n_bd * pitch ** 3, grid_shape=shape,
               n_fully_inside=n_in, n_boundary=n_bd, n_fine_inside=0, n_fine_boundary=0,
               n_points=int(np.prod(shape)))
    if n_bd and r > 1:
        fine = pitch / r
        f_in, f_bd = passes.refine(lo, pitch, shape, n_bd, r, 0.5 * fine * math.sqrt(3.0))
        out.update(volume_mm3=out["volume_mm3"] + f_in * fine ** 3, uncertainty_mm3=f_bd * fine ** 3,
                   n_fine_inside=f_in, n_fine_boundary=f_bd, n_points=out["n_points"] + n_bd * r ** 3)
    out.update(coarse_pitch_mm=pitch, refine_factor=r, fine_pitch_mm=pitch / r, empty_overlap=False)
    return out
