"""Hierarchical volume queries over IKARUS trees (FE/ikarus_v1/queries.py:36-277,
FE/ikarus_v1/cost_model_v1_2.py).

`intrusion_volume(a, b, budget_ms)` keeps the reference's method and result
dictionary: a two-term cost model (fixed overhead + points per second) sizes a
probe pass, a coarse pass over the overlap of the operands' boxes and a
refinement of the coarse cells the Lipschitz bound cannot decide; the volume
is (coarse cells with v < -half diagonal) + (fine samples with v < 0), and
the declared uncertainty is the volume of the cells still undecided at the
finest level.

Here every pass is one device kernel over an implicit grid (csrc/ikarus.hip
`k_ik_grid`, `k_ik_refine`): samples are generated on the device from
(lo, pitch, shape) with the same f64 arithmetic as NumPy's grid and rounded
to fp32 as eval_batch would, classified in registers and counted with
block-level reductions; only the undecided cells' indices are kept (for the
refinement), so no point array crosses the bus.  `grid_volume` is the
deterministic core for given pitches; `intrusion_volume` adds the budget
dispatch.
"""
import hashlib
import json
import math
import os
import time

import numpy as np

from . import _lib
from . import ikarus as E

_COST_CACHE = {}


def bbox_of(node, pad=1.0):
    """Axis-aligned box of a tree from its leaves' declared extents
    (FE/ikarus_v1/queries.py:36-95): sphere axis points, box and
    cylinder/cone corners, a grid leaf's sample extent; transforms map the
    points, offset/scale/axis_project pass the child's box through unchanged,
    halfspaces contribute nothing."""
    lo = np.array([1e30, 1e30, 1e30])
    hi = -lo.copy()

    def walk(n, xform):
        nonlocal lo, hi
        op = n["op"]
        if op == "sphere":
            c = np.array(n["center"])
            r = n["radius"]
            pts = np.array([c + r * np.array(v) for v in
                            [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]])
        elif op == "box":
            c = np.array(n["center"])
            h = np.array(n["half_extents"])
            pts = np.array([c + np.array([sx * h[0], sy * h[1], sz * h[2]])
                            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        elif op in ("cylinder", "cone"):
            c = np.array(n["center"])
            h = n["height"]
            r = max(n.get("radius", 0), n.get("radius1", 0), n.get("radius2", 0))
            pts = np.array([c + np.array([sx * r, sy * r, sz * h / 2.0])
                            for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        elif op == "rect_frustum":
            hx, hy = max(n["hx0"], n["hx1"]), max(n["hy0"], n["hy1"])
            cxs, cys = (n["cx0"], n["cx1"]), (n["cy0"], n["cy1"])
            pts = np.array([[cx + sx * hx, cy + sy * hy, z] for cx in cxs for cy in cys
                            for sx in (-1, 1) for sy in (-1, 1) for z in (n["z0"], n["z1"])])
        elif op == "halfspace":
            return
        elif op == "gridfield":
            arr, origin, pitch, _, _ = E._grid_samples(n)
            origin = np.asarray(origin, np.float64)
            pts = np.array([origin, origin + pitch * (np.array(arr.shape) - 1)])
        elif op == "transform":
            walk(n["child"], xform @ np.array(n["mat4"]))
            return
        elif op in ("offset", "scale", "axis_project"):
            walk(n["child"], xform)
            return
        elif op in ("union", "intersect", "subtract"):
            walk(n["a"], xform)
            walk(n["b"], xform)
            return
        else:
            raise ValueError(op)
        world = (xform @ np.hstack([pts, np.ones((len(pts), 1))]).T).T[:, :3]
        lo = np.minimum(lo, world.min(axis=0))
        hi = np.maximum(hi, world.max(axis=0))

    walk(node, np.eye(4))
    return lo - pad, hi + pad


def grid_shape(lo, hi, pitch):
    """Samples per axis of the grid lo + pitch*i covering [lo, hi] (FE/ikarus_v1/queries.py:120-131)."""
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
    FE/ikarus_v1/queries.py:211-264).  Returns a dict with the counts and the
    declared uncertainty."""
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


def calibrate_two_term(key, passes, lo, hi, n_lo=20_000, n_hi=400_000, reps=5, cache_path=None,
                       force_remeasure=False):
    """(overhead_s, rate_pts_per_s, from_cache) of one grid pass, fitted from
    two sizes by the line through their median timings
    (FE/ikarus_v1/cost_model_v1_2.py:64-104); cached per key in memory and, when
    cache_path is given, in a JSON file."""
    if not force_remeasure:
        if key in _COST_CACHE:
            return (*_COST_CACHE[key], True)
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    e = json.load(f).get(key)
                if e:
                    _COST_CACHE[key] = (e["overhead_s"], e["rate_pts_per_s"])
                    return e["overhead_s"], e["rate_pts_per_s"], True
            except (OSError, json.JSONDecodeError):
                pass

    def timed(n):
        pitch = max((float(np.prod(np.asarray(hi) - np.asarray(lo))) / n) ** (1.0 / 3.0), 1e-6)
        shape = grid_shape(lo, hi, pitch)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            passes.grid(lo, pitch, shape, 0.0, 0.0, keep_cells=False)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), int(np.prod(shape))

    t_lo, m_lo = timed(n_lo)
    t_hi, m_hi = timed(n_hi)
    dn, dt = m_hi - m_lo, t_hi - t_lo
    if dn <= 0 or dt <= 0:
        rate, overhead = m_hi / max(t_hi, 1e-6), 0.0
    else:
        rate = dn / dt
        overhead = max(t_lo - m_lo / rate, 0.0)
    _COST_CACHE[key] = (overhead, rate)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    cache = json.load(f)
            except (OSError, json.JSONDecodeError):
                cache = {}
        cache[key] = {"overhead_s": overhead, "rate_pts_per_s": rate, "n_lo": m_lo, "t_lo_s": t_lo, "n_hi": m_hi,
                      "t_hi_s": t_hi, "reps": reps, "measured_at_unix": time.time()}
        tmp = str(cache_path) + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=2, sort_keys=True)
        os.replace(tmp, cache_path)
    return overhead, rate, False


def affordable_points(budget_ms, overhead_s, rate_pts_per_s, floor_points=64):
    """Points a budget affords under the two-term model (FE/ikarus_v1/cost_model_v1_2.py:106-114)."""
    remaining_s = budget_ms / 1000.0 - overhead_s
    if remaining_s <= 0:
        return floor_points
    return max(int(remaining_s * rate_pts_per_s), floor_points)


def intrusion_volume(exprA, exprB, budget_ms=200.0, min_pitch_mm=0.25, max_refine_factor=8, device="cuda:0",
                     log=None, max_coarse_points=None, cache_path=None):
    """Volume of A intersect B with a declared uncertainty, sized by a compute
    budget (FE/ikarus_v1/queries.py:138-277): probe pass for the undecided
    fraction, coarse pass, refinement of undecided coarse cells."""
    t_start = time.perf_counter()
    combo = E.intersect(exprA, exprB, k=0.0)
    passes = _Passes(combo, device)
    loA, hiA = bbox_of(exprA)
    loB, hiB = bbox_of(exprB)
    lo = np.maximum(loA, loB)
    hi = np.minimum(hiA, hiB)
    if np.any(hi <= lo):
        return {"method": "hierarchical_prune_adaptive_refine", "volume_mm3": 0.0, "uncertainty_mm3": 0.0,
                "budget_ms": budget_ms, "wall_ms": 0.0, "budget_met_within_20pct": True,
                "dispatch_log": {"note": "operand bboxes do not overlap -- 0 volume by construction, no grid pass "
                                         "needed", "bboxA": [loA.tolist(), hiA.tolist()],
                                 "bboxB": [loB.tolist(), hiB.tolist()]}}
    key = "ikarus_grid::" + hashlib.sha256(passes.prog.ins[:, :2].tobytes()).hexdigest()[:16] + "::" + str(passes.dev)
    overhead_s, rate, from_cache = calibrate_two_term(key, passes, lo, hi, cache_path=cache_path)
    budget_points = affordable_points(budget_ms, overhead_s, rate, floor_points=1000)
    bbox_vol = float(np.prod(hi - lo))

    t0 = time.perf_counter()
    probe_pitch = max((bbox_vol / max(int(0.03 * budget_points), 200)) ** (1.0 / 3.0), min_pitch_mm)
    probe_shape = grid_shape(lo, hi, probe_pitch)
    n_probe = int(np.prod(probe_shape))
    probe_diag_half = 0.5 * probe_pitch * math.sqrt(3.0)
    _, probe_bd = passes.grid(lo, probe_pitch, probe_shape, -probe_diag_half, probe_diag_half, keep_cells=False)
    t_probe = time.perf_counter() - t0
    f_boundary = min(max(probe_bd / n_probe, 1e-4), 1.0)

    target_refine_factor = 4
    remaining_after_probe = max(budget_points - n_probe, 1000)
    denom = 1.0 + f_boundary * (target_refine_factor ** 3)
    n_coarse_target = max(int(remaining_after_probe / denom), 64)
    if max_coarse_points is not None:
        n_coarse_target = min(n_coarse_target, max_coarse_points)
    coarse_pitch = max((bbox_vol / n_coarse_target) ** (1.0 / 3.0), min_pitch_mm)

    t0 = time.perf_counter()
    shape_c = grid_shape(lo, hi, coarse_pitch)
    n_coarse = int(np.prod(shape_c))
    coarse_diag_half = 0.5 * coarse_pitch * math.sqrt(3.0)
    n_inside, n_boundary = passes.grid(lo, coarse_pitch, shape_c, -coarse_diag_half, coarse_diag_half,
                                       keep_cells=True)
    t_coarse = time.perf_counter() - t0
    vol_from_coarse = float(n_inside) * coarse_pitch ** 3

    remaining_points = max(budget_points - n_probe - n_coarse, 0)
    refine_factor = min(max_refine_factor, target_refine_factor)
    if n_boundary > 0:
        while refine_factor > 1 and n_boundary * (refine_factor ** 3) > max(remaining_points, 1):
            refine_factor -= 1
    refine_factor = max(refine_factor, 1)
    fine_pitch = max(coarse_pitch / refine_factor, min_pitch_mm)

    t0 = time.perf_counter()
    vol_from_fine = 0.0
    n_fine_pts = n_boundary_fine = 0
    if n_boundary > 0 and refine_factor > 1:
        fine_diag_half = 0.5 * (coarse_pitch / refine_factor) * math.sqrt(3.0)
        fine_in, n_boundary_fine = passes.refine(lo, coarse_pitch, shape_c, n_boundary, refine_factor,
                                                 fine_diag_half)
        n_fine_pts = n_boundary * refine_factor ** 3
        vol_from_fine = float(fine_in) * (coarse_pitch / refine_factor) ** 3
    t_fine = time.perf_counter() - t0

    if n_boundary > 0 and refine_factor > 1:
        uncertainty = n_boundary_fine * (coarse_pitch / refine_factor) ** 3
    else:
        uncertainty = n_boundary * coarse_pitch ** 3
    wall_ms = (time.perf_counter() - t_start) * 1000.0
    result = {
        "method": "hierarchical_prune_adaptive_refine",
        "volume_mm3": vol_from_coarse + vol_from_fine,
        "uncertainty_mm3": uncertainty,
        "budget_ms": budget_ms,
        "wall_ms": wall_ms,
        "budget_met_within_20pct": bool(wall_ms <= budget_ms * 1.2),
        "dispatch_log": {
            "calibrated_rate_pts_per_s": rate,
            "cost_model_v1_2": {"overhead_s": overhead_s, "rate_pts_per_s": rate, "from_cache": from_cache},
            "budget_points": budget_points, "max_coarse_points_cap": max_coarse_points,
            "coarse_cap_engaged": bool(max_coarse_points is not None and n_coarse_target >= max_coarse_points),
            "probe_pitch_mm": probe_pitch, "n_probe_points": n_probe, "probe_wall_s": t_probe,
            "boundary_fraction_estimate": f_boundary, "target_refine_factor": target_refine_factor,
            "bbox_mm": {"lo": lo.tolist(), "hi": hi.tolist(), "volume_mm3": bbox_vol},
            "coarse_pitch_mm": coarse_pitch, "coarse_grid_shape": shape_c, "n_coarse_points": n_coarse,
            "n_fully_inside_coarse": n_inside, "n_boundary_coarse": n_boundary, "coarse_wall_s": t_coarse,
            "refine_factor_chosen": refine_factor, "fine_pitch_mm": fine_pitch, "n_fine_points": n_fine_pts,
            "n_boundary_remaining_at_finest": n_boundary_fine, "fine_wall_s": t_fine,
            "space_eliminated_frac": 1.0 - (n_boundary / max(n_coarse, 1)),
        },
    }
    if log is not None:
        log.append(result["dispatch_log"])
    return result
