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


def predicted_wall_ms(n_points, overhead_s, rate_pts_per_s):
    """FE/ikarus_v1/cost_model_v1_2.py:111-112."""
    return (overhead_s + n_points / max(rate_pts_per_s, 1e-9)) * 1000.0


def total_volume(node, budget_ms=200.0, min_pitch_mm=0.25, device="cuda:0"):
    """Volume of one tree: intrusion_volume(node, node) (FE/ikarus_v1/queries.py:554-558)."""
    return intrusion_volume(node, node, budget_ms=budget_ms, min_pitch_mm=min_pitch_mm, device=device)


def calibrate_throughput(node, n_probe=200_000, device="cuda:0"):
    """(points per second, seconds) of one eval_batch call over n_probe seeded
    uniform points in the tree's box (FE/ikarus_v1/queries.py:96-106)."""
    lo, hi = bbox_of(node)
    pts = np.random.default_rng(0).uniform(lo, hi, size=(n_probe, 3))
    t0 = time.perf_counter()
    E.eval_batch(node, pts, device=device)
    dt = time.perf_counter() - t0
    return n_probe / max(dt, 1e-6), dt


def calibrate_throughput_cached(node, device="cuda:0", cache_path=None):
    """(rate, overhead_s, from_cache) of the two-term model of a tree's grid
    pass, cached per program shape (FE/ikarus_v1/queries.py:109-122)."""
    passes = _Passes(node, device)
    lo, hi = bbox_of(node)
    key = "ikarus_grid::" + hashlib.sha256(passes.prog.ins[:, :2].tobytes()).hexdigest()[:16] + "::" + str(passes.dev)
    overhead_s, rate, from_cache = calibrate_two_term(key, passes, lo, hi, cache_path=cache_path)
    return rate, overhead_s, from_cache


def min_clearance(exprA, exprB, budget_ms=200.0, n_samples=None, device="cuda:0"):
    """Clearance of B from A's surface (FE/ikarus_v1/queries.py:561-597): A's
    samples within half a cell diagonal of its surface on a budget-sized grid,
    then the minimum of B's value there.  The near-surface samples are kept on
    the device (the grid pass's band list) and B is evaluated at them there."""
    import torch
    rate, _ = calibrate_throughput(exprA, n_probe=50_000, device=device)
    budget_points = max(int(budget_ms / 1000.0 * rate * 0.5), 500)
    lo, hi = bbox_of(exprA)
    bbox_vol = float(np.prod(hi - lo))
    pitch = max((bbox_vol / budget_points) ** (1.0 / 3.0), 0.25)
    t0 = time.perf_counter()
    shape = grid_shape(lo, hi, pitch)
    diag_half = 0.5 * pitch * math.sqrt(3.0)
    passes = _Passes(exprA, device)
    _, n_surf = passes.grid(lo, pitch, shape, -math.inf, diag_half, keep_cells=True)
    if n_surf == 0:
        return {"method": "surface_sample_vs_field", "min_clearance_mm": float("inf"), "n_surface_samples": 0,
                "wall_ms": (time.perf_counter() - t0) * 1000.0,
                "note": "expr A empty at this bbox/pitch -- no surface samples found"}
    idx = passes.cells[:n_surf].to(torch.int64)
    ny, nz = shape[1], shape[2]
    ijk = torch.stack([idx // (ny * nz), (idx // nz) % ny, idx % nz], 1).to(torch.float64)
    lo_t = torch.tensor(np.asarray(lo, np.float64), device=idx.device)
    # lo + pitch * i in f64 (two roundings, as NumPy's grid), then fp32 as eval_batch casts
    pts = (lo_t + ijk * pitch).to(torch.float32).contiguous()
    vals = torch.empty(n_surf, dtype=torch.float32, device=idx.device)
    E.eval_device(E.compile_tree(exprB), pts, vals)
    min_clear = float(vals.min().item())
    wall_s = time.perf_counter() - t0
    return {"method": "surface_sample_vs_field", "min_clearance_mm": min_clear, "n_surface_samples": int(n_surf),
            "pitch_mm": pitch, "uncertainty_mm": pitch, "budget_ms": budget_ms, "wall_ms": wall_s * 1000.0,
            "budget_met_within_20pct": bool(wall_s * 1000.0 <= budget_ms * 1.2)}


# ------------------------------------------------------------------ batched pair sweep
PAIR_REC = np.dtype([("lx", "<f8"), ("ly", "<f8"), ("lz", "<f8"), ("pitch", "<f8"), ("first", "<i8"),
                     ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("la", "<i4"), ("lb", "<i4"),
                     ("below", "<f4"), ("band", "<f4"), ("fband", "<f4"), ("pad", "<i4", (2,))])
assert PAIR_REC.itemsize == 80


class _PairPasses:
    """Device state of a batched pair sweep: the leaf table, one pair table
    per pass, per-pair counters and the band-sample list."""

    def __init__(self, leaves, device):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.nleaves = len(leaves)
        self.grids, self.gdata = E.leaf_table(leaves, self.dev)
        self.cells = None
        self.cursor = torch.zeros(1, dtype=torch.int32, device=self.dev)

    def upload(self, recs):
        return self.torch.from_numpy(np.ascontiguousarray(recs).view(np.uint8)).to(self.dev)

    def coarse(self, recs):
        """Counts [n, 2] (below, band) of the coarse pass over the pairs `recs`;
        the band samples' global indices stay in self.cells."""
        t = self.torch
        P = _lib.ptr
        total = int(recs["first"][-1]) + int(recs["nx"][-1]) * int(recs["ny"][-1]) * int(recs["nz"][-1])
        if total >= 2 ** 31 - 1:
            raise ValueError("a sweep pass holds at most 2^31 - 2 samples")
        counts = t.zeros((len(recs), 2), dtype=t.int64, device=self.dev)
        if self.cells is None or self.cells.numel() < max(total, 1):
            self.cells = t.empty(max(total, 1), dtype=t.int32, device=self.dev)
        self.cursor.zero_()
        self.table = self.upload(recs)
        E._check(_lib.lib().fe_ikarus_pairs_grid(P(self.table), len(recs), total, P(self.grids), self.nleaves,
                                                 P(self.gdata), P(counts), P(self.cells), P(self.cursor),
                                                 _lib.stream_handle(None)), "pair sweep coarse pass")
        return counts.cpu().numpy()

    def refine(self, recs, ncells, r):
        """Counts [n, 2] (below 0, fine band) of r^3 samples per band sample of
        the last coarse pass (recs: the same pairs with the fine bands set)."""
        t = self.torch
        P = _lib.ptr
        counts = t.zeros((len(recs), 2), dtype=t.int64, device=self.dev)
        table = self.upload(recs)
        E._check(_lib.lib().fe_ikarus_pairs_refine(P(table), len(recs), P(self.grids), self.nleaves, P(self.gdata),
                                                   P(self.cells), int(ncells), int(r), P(counts),
                                                   _lib.stream_handle(None)), "pair sweep refine pass")
        return counts.cpu().numpy()


class _OnePair:
    """The coarse pass of one (leaf, leaf) pair as a `grid` call, for the cost model."""

    def __init__(self, passes, la, lb):
        self.p, self.la, self.lb = passes, la, lb

    def grid(self, lo, pitch, shape, below, band, keep_cells):
        rec = np.zeros(1, PAIR_REC)
        rec["lx"], rec["ly"], rec["lz"] = lo
        rec["pitch"] = pitch
        rec["nx"], rec["ny"], rec["nz"] = shape
        rec["la"], rec["lb"] = self.la, self.lb
        rec["below"], rec["band"] = np.float32(below), np.float32(band)
        c = self.p.coarse(rec)
        return int(c[0, 0]), int(c[0, 1])


def scene_sweep_gridfield_pairs(pairs, total_budget_ms=800.0, min_pitch_mm=0.25, max_refine_factor=8,
                                target_refine_factor=4, min_pair_coarse_points=27, device="cuda:0", cache_path=None,
                                prior_vol_override=None, max_pair_frac_of_pool=0.05, max_pair_points_abs=150_000,
                                max_total_refine_points_abs=6_000_000):
    """Hard-intersection volume and uncertainty of many (grid leaf, grid leaf)
    pairs under one sweep budget (FE/ikarus_v1/queries.py:281-551): the budget
    is shared by the pairs' box-overlap priors (clipped per pair), every pair's
    coarse grid runs in ONE device launch and every pair's band samples are
    refined in a second.  Returns (results, sweep_log) with the reference's
    keys.  The grids are generated on the device from (lo, pitch, shape), so
    no point buffer is built on the host."""
    n_pairs = len(pairs)
    if n_pairs == 0:
        return [], {"n_pairs": 0, "note": "empty pair list"}
    t_sweep0 = time.perf_counter()

    # the distinct leaves (the device leaf table) and each pair's two of them
    all_leaves, leaf_index = [], {}

    def leaf_id(node):
        k = E._leaf_key(node)
        if k not in leaf_index:
            leaf_index[k] = len(all_leaves)
            all_leaves.append(node)
        return leaf_index[k]

    leaf_a_of_pair = np.array([leaf_id(la) for la, _ in pairs], np.int32)
    leaf_b_of_pair = np.array([leaf_id(lb) for _, lb in pairs], np.int32)

    # stage A: box-overlap priors, over arrays of pairs (one box per distinct
    # leaf; the volume is (dx*dy)*dz, as a product over the three extents)
    boxes = [bbox_of(leaf) for leaf in all_leaves]
    leaf_lo = np.array([b[0] for b in boxes], np.float64)
    leaf_hi = np.array([b[1] for b in boxes], np.float64)
    pr_lo = np.maximum(leaf_lo[leaf_a_of_pair], leaf_lo[leaf_b_of_pair])
    pr_hi = np.minimum(leaf_hi[leaf_a_of_pair], leaf_hi[leaf_b_of_pair])
    pr_empty = np.any(pr_hi <= pr_lo, axis=1)
    ext = pr_hi - pr_lo
    pr_vol = np.where(pr_empty, 0.0, ext[:, 0] * ext[:, 1] * ext[:, 2])
    if prior_vol_override is not None:
        if len(prior_vol_override) != n_pairs:
            raise ValueError("prior_vol_override needs one value per pair")
        alloc_vol = [0.0 if e else float(v) for e, v in zip(pr_empty.tolist(), prior_vol_override)]
    else:
        alloc_vol = pr_vol.tolist()
    vol_list, empty_list = pr_vol.tolist(), pr_empty.tolist()
    total_prior_vol = sum(alloc_vol)  # left to right, like the per-pair loop
    t_prior = time.perf_counter() - t_sweep0
    passes = _PairPasses(all_leaves, device)

    # stage B: two-term cost model of the pair pass (cached per device)
    lo0, hi0 = boxes[int(leaf_a_of_pair[0])]
    overhead_s, rate, from_cache = calibrate_two_term("ikarus_pairs::" + str(passes.dev),
                                                      _OnePair(passes, int(leaf_a_of_pair[0]),
                                                               int(leaf_a_of_pair[0])),
                                                      lo0, hi0, cache_path=cache_path)
    floor_total = n_pairs * min_pair_coarse_points
    total_points_budget = affordable_points(total_budget_ms, overhead_s, rate, floor_points=floor_total)
    alloc_pool = max(total_points_budget - floor_total, 0)

    # stage C: per-pair coarse pitch from its clipped share of the pool
    max_pair_points = max(min(int(alloc_pool * max_pair_frac_of_pool), max_pair_points_abs), min_pair_coarse_points)
    coarse_pitch_of_pair = [None] * n_pairs
    n_pairs_capped = 0
    for i in range(n_pairs):
        if empty_list[i]:
            continue
        share = (alloc_vol[i] / total_prior_vol) if total_prior_vol > 0 else (1.0 / n_pairs)
        n_target = min_pair_coarse_points + int(alloc_pool * share)
        if n_target > max_pair_points:
            n_target = max_pair_points
            n_pairs_capped += 1
        coarse_pitch_of_pair[i] = max((vol_list[i] / max(n_target, 1)) ** (1.0 / 3.0), min_pitch_mm)

    # the pair table of the coarse launch (pairs with a grid pass, in order),
    # filled column by column: grid_shape's ceil((hi - lo) / pitch) + 1 (at
    # least 1) per axis, band half-width 0.5 * pitch * sqrt(3)
    t_buf0 = time.perf_counter()
    live = [i for i in range(n_pairs) if coarse_pitch_of_pair[i] is not None]
    live_ix = np.array(live, np.int64)
    recs = np.zeros(len(live), PAIR_REC)
    first = 0
    if len(live):
        pitch_arr = np.array([coarse_pitch_of_pair[i] for i in live], np.float64)
        lo_l, hi_l = pr_lo[live_ix], pr_hi[live_ix]
        shape_arr = np.maximum(np.ceil((hi_l - lo_l) / pitch_arr[:, None]).astype(np.int64) + 1, 1)
        npts = shape_arr[:, 0] * shape_arr[:, 1] * shape_arr[:, 2]
        dh = 0.5 * pitch_arr * math.sqrt(3.0)
        recs["lx"], recs["ly"], recs["lz"] = lo_l[:, 0], lo_l[:, 1], lo_l[:, 2]
        recs["pitch"] = pitch_arr
        recs["first"] = np.concatenate(([0], np.cumsum(npts)[:-1]))
        recs["nx"], recs["ny"], recs["nz"] = shape_arr[:, 0], shape_arr[:, 1], shape_arr[:, 2]
        recs["la"], recs["lb"] = leaf_a_of_pair[live_ix], leaf_b_of_pair[live_ix]
        recs["below"], recs["band"] = (-dh).astype(np.float32), dh.astype(np.float32)
        first = int(npts.sum())
    t_buffer_build = time.perf_counter() - t_buf0

    t0 = time.perf_counter()
    coarse = passes.coarse(recs) if len(live) else np.zeros((0, 2), np.int64)
    t_coarse_launch = time.perf_counter() - t0

    # stage D: per-pair volume and band counts of the coarse launch
    results = [None] * n_pairs
    n_boundary_of_pair = [0] * n_pairs
    for i in range(n_pairs):
        if coarse_pitch_of_pair[i] is None:
            results[i] = {"method": "sweep_batched_prior_alloc", "volume_mm3": 0.0, "uncertainty_mm3": 0.0,
                          "dispatch_log": {"note": "AABB-overlap prior volume is 0 -- no grid pass needed",
                                           "prior_vol_mm3": 0.0}}
    npts_list = npts.tolist() if len(live) else []
    coarse_list = coarse.tolist()
    for r_i, i in enumerate(live):
        pitch = coarse_pitch_of_pair[i]
        n_i = npts_list[r_i]
        n_in, n_b = coarse_list[r_i]
        cell_vol = pitch ** 3
        n_boundary_of_pair[i] = n_b
        results[i] = {"method": "sweep_batched_prior_alloc", "volume_mm3": float(n_in) * cell_vol,
                      "uncertainty_mm3": n_b * cell_vol,
                      "dispatch_log": {"prior_vol_mm3": vol_list[i], "coarse_pitch_mm": pitch,
                                       "n_coarse_points": n_i, "n_fully_inside_coarse": n_in,
                                       "n_boundary_coarse": n_b}}

    # stage E: one refinement launch over every pair's band samples
    total_boundary = sum(n_boundary_of_pair)
    remaining_budget_ms = max(total_budget_ms - t_coarse_launch * 1000.0, 0.0)
    remaining_points_budget = affordable_points(remaining_budget_ms, overhead_s, rate, floor_points=0)
    refine_factor = 1
    t_refine_launch = 0.0
    t_refine_buffer_build = 0.0
    if total_boundary > 0 and remaining_points_budget > total_boundary:
        refine_factor = int(min(max_refine_factor, target_refine_factor,
                                max((remaining_points_budget / total_boundary) ** (1.0 / 3.0), 1.0)))
    while refine_factor > 1 and total_boundary * (refine_factor ** 3) > max_total_refine_points_abs:
        refine_factor -= 1
    if refine_factor > 1 and total_boundary > 0:
        t_rb0 = time.perf_counter()
        recs["fband"] = (0.5 * (pitch_arr / refine_factor) * math.sqrt(3.0)).astype(np.float32)
        t_refine_buffer_build = time.perf_counter() - t_rb0
        t0 = time.perf_counter()
        fine = passes.refine(recs, total_boundary, refine_factor)
        t_refine_launch = time.perf_counter() - t0
        fine_list = fine.tolist()
        for r_i, i in enumerate(live):
            if n_boundary_of_pair[i] == 0:
                continue
            raw_fp = coarse_pitch_of_pair[i] / refine_factor
            fine_cell_vol = raw_fp ** 3
            results[i]["volume_mm3"] += float(int(fine_list[r_i][0])) * fine_cell_vol
            results[i]["uncertainty_mm3"] = int(fine_list[r_i][1]) * fine_cell_vol
            log = results[i]["dispatch_log"]
            log["refine_factor"] = refine_factor
            log["fine_pitch_mm"] = max(raw_fp, min_pitch_mm)
            log["n_fine_points"] = n_boundary_of_pair[i] * refine_factor ** 3
            log["n_boundary_remaining_at_finest"] = int(fine_list[r_i][1])

    sweep_log = {
        "n_pairs": n_pairs, "total_budget_ms": total_budget_ms,
        "cost_model_v1_2": {"overhead_s": overhead_s, "rate_pts_per_s": rate, "from_cache": from_cache},
        "total_prior_overlap_vol_mm3": total_prior_vol,
        "n_pairs_empty_prior": int(np.count_nonzero(pr_empty)),
        "total_points_budget": total_points_budget, "n_coarse_points_actual": int(first),
        "max_pair_points_cap": max_pair_points, "n_pairs_capped": n_pairs_capped,
        "n_launches": 1 + (1 if (refine_factor > 1 and total_boundary > 0) else 0),
        "refine_factor_chosen": refine_factor, "total_boundary_cells": total_boundary,
        "t_prior_stage_s": t_prior, "t_buffer_build_s": t_buffer_build, "t_coarse_launch_s": t_coarse_launch,
        "t_refine_buffer_build_s": t_refine_buffer_build, "t_refine_launch_s": t_refine_launch,
        "t_sweep_wall_s": time.perf_counter() - t_sweep0,
    }
    return results, sweep_log
