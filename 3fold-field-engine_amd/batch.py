"""Batched builds: many uneven recipes in one captured HIP graph launch.

A batch is split by block edge into FieldPlans (one descriptor table each);
the classify, compaction and persistent tile launches of every plan are
captured once into a HIP graph (through torch.cuda.CUDAGraph), so a whole
batch of builds is one graph launch with no host synchronisation.  The
arithmetic is the plain path's: every field equals its per-call build with the
same block edge bit for bit.

The block edge (4, 8 or 16) and the number of tile waves of the persistent
tile launch are picked per domain-size class by `Autotuner`, which times
candidate configurations on a sample of the class on the device.
"""
import time

import numpy as np

from . import field as F
from . import recipes as R

SIZE_CLASSES = ((0, 128), (129, 256), (257, 1 << 20))
BLOCKS = (4, 8, 16)
TILE_WAVES = (1024, 4096, 16384)


def _torch():
    import torch
    return torch


def size_class(spec):
    n = max(spec[4], spec[5], spec[6])
    for i, (lo, hi) in enumerate(SIZE_CLASSES):
        if lo <= n <= hi:
            return i
    return len(SIZE_CLASSES) - 1


def make_meta(spec, block):
    """spec = (x0, y0, z0, pitch, nx, ny, nz)."""
    return F.GridMeta(*spec[:4], *spec[4:7], block=block)


def _event_ms(fn, reps=3):
    torch = _torch()
    fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


class Autotuner:
    """Chooses (block edge, tile waves) per domain-size class by timing builds of
    up to `sample` recipes of the class with every candidate on the device."""

    def __init__(self, blocks=BLOCKS, tile_waves=TILE_WAVES, sample=6, device=None):
        self.blocks, self.tile_waves, self.sample = tuple(blocks), tuple(tile_waves), sample
        self.device = device
        self.table = {}
        self.timings = {}

    def tune(self, items):
        """items: [(spec, recipe)].  Fills and returns {class: (block, waves)}."""
        by_class = {}
        for spec, rec in items:
            by_class.setdefault(size_class(spec), []).append((spec, rec))
        for cls, members in sorted(by_class.items()):
            if cls in self.table:
                continue
            sample = members[:self.sample]
            best = None
            for B in self.blocks:
                plan = F.FieldPlan([(make_meta(s, B), r) for s, r in sample], device=self.device)
                for w in self.tile_waves:
                    plan.tile_waves = w
                    ms = _event_ms(plan.launch)
                    self.timings[(cls, B, w)] = ms
                    if best is None or ms < best[0]:
                        best = (ms, B, w)
                del plan
            self.table[cls] = (best[1], best[2])
        return self.table

    def choose(self, spec):
        return self.table.get(size_class(spec), (8, 4096))


class BatchBuilder:
    """Builds a list of recipes in one graph launch.

    items: [(spec, recipe)] with spec = (x0, y0, z0, pitch, nx, ny, nz).
    block: a fixed block edge for every item, or None to use `tuner` (a tuned
    Autotuner; a fresh one is tuned on the items when omitted)."""

    def __init__(self, items, block=None, tuner=None, device=None, capture=True):
        torch = _torch()
        self.device = torch.device(device or "cuda:0")
        if block is None:
            if tuner is None:
                tuner = Autotuner(device=self.device)
                tuner.tune(items)
            choice = [tuner.choose(spec) for spec, _ in items]
        else:
            choice = [(block, 4096)] * len(items)
        self.choice = choice
        groups = {}
        for i, ((spec, rec), (B, w)) in enumerate(zip(items, choice)):
            groups.setdefault(B, []).append(i)
        self.plans, self.where = [], [None] * len(items)
        for B, idx in sorted(groups.items()):
            plan = F.FieldPlan([(make_meta(items[i][0], B), items[i][1]) for i in idx], device=self.device)
            plan.tile_waves = max(choice[i][1] for i in idx)
            for j, i in enumerate(idx):
                self.where[i] = (len(self.plans), j)
            self.plans.append(plan)
        self.graph = None
        if capture:
            self._capture()

    def _capture(self):
        torch = _torch()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch()  # warm-up outside the graph (lazy module loading)
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch()
        torch.cuda.synchronize(self.device)

    def _launch(self):
        for plan in self.plans:
            plan.launch()

    def run(self):
        """One batch: a graph replay (or the eager launches without capture)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch()

    def __len__(self):
        return len(self.where)

    def plan_of(self, i):
        p, j = self.where[i]
        return self.plans[p], j

    def sparse_field(self, i):
        plan, j = self.plan_of(i)
        return plan.sparse_field(j)

    def n_active(self, i):
        plan, j = self.plan_of(i)
        return plan.n_active(j)


def build_batch(items, block=None, tuner=None, device=None):
    """Builds every (spec, recipe) in one graph launch; returns the BatchBuilder
    (results stay resident; sparse_field(i) downloads one)."""
    b = BatchBuilder(items, block=block, tuner=tuner, device=device)
    b.run()
    return b


def plain_builds(items, block=8, device="cuda:0"):
    """The plain per-call path for each item (host compaction, per-call allocation)."""
    return [F.klassificera_och_evaluera_recept(make_meta(spec, block), rec, device=device) for spec, rec in items]


def batch_workload(seed=3, count=256):
    """Seeded uneven recipes: [(spec, recipe)] with cubic domains at 1 mm."""
    return [((0.0, 0.0, 0.0, 1.0, n, n, n), rec) for rec, n in R.batch_recipes(seed, count)]


def time_batch_vs_plain(items, reps=3, plain_items=None, device="cuda:0"):
    """fields/s of the graph-captured batch and of sequential plain per-call builds
    (each plain build uses the block edge the batch chose for that item)."""
    torch = _torch()
    tuner = Autotuner(device=device)
    tuner.tune(items)
    b = BatchBuilder(items, tuner=tuner, device=device)
    b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / reps
    sub = plain_items if plain_items is not None else list(range(len(items)))
    t1 = time.perf_counter()
    for i in sub:
        spec, rec = items[i]
        F.klassificera_och_evaluera_recept(make_meta(spec, b.choice[i][0]), rec, device=device)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t1) * len(items) / len(sub)
    return dict(batch_fields_s=len(items) / batch_s, plain_fields_s=len(items) / plain_s,
                batch_ms=batch_s * 1e3, plain_ms=plain_s * 1e3, speedup=plain_s / batch_s,
                tuning={f"class{k}": v for k, v in tuner.table.items()}), b
