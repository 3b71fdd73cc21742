"""Field server: continuous batching of field builds under an HBM budget, with
built fields kept resident on the device for point-SDF queries.

Requests (grid spec + recipe) queue on submit.  A worker thread that owns the
device streams admits queued requests in arrival order while (a) a batch slot
(one HIP stream) is free and (b) the transient build workspace plus every
resident field fits the HBM budget.  Each admitted batch is one FieldPlan
launch on its own stream, so a batch admitted while another is still on the
GPU overlaps with it.  When a batch's completion event fires, every field is
compacted to its resident form (kinds, active ids, tiles, culled lists and
its own descriptor), the workspace is released and the request completes.

The build arithmetic is the plain path's: a field built in a mixed batch is
bit-identical to the same recipe built alone.  Latencies are host wall-clock:
build = arrival -> resident; query = call -> result.
"""
import collections
import itertools
import threading
import time

import numpy as np

from . import _lib
from . import field as F
from . import recipes as R


def _torch():
    import torch
    return torch


SEG = np.dtype([("src", "<u8"), ("dst", "<u8"), ("n", "<i8")])  # fe_copy_segments record


class ResidentField:
    """A built field kept on the device: SparseField data plus a descriptor,
    all in one device buffer (`store`; the arrays are views into it).
    n_active: the build's active-block count when the caller has read it
    already.  copy=False / upload=False leave the copies (`segments`) and the
    descriptor upload to the caller, which batches them (resident_fields)."""

    _PIECES = (("kind", 4), ("ids", 4), ("tiles", 4), ("cma", 4), ("slot", 4), ("list_len", 4), ("lists", 2),
               ("prims", 1), ("counts", 4))

    def __init__(self, plan, j, stream, n_active=None, upload=True, copy=True):
        torch = _torch()
        meta, rec = plan.items[j]
        self.meta, self.recipe = meta, rec
        na = plan.n_active(j) if n_active is None else n_active
        self.n_active = na
        self.B = plan.B
        off = sum(len(r) for _, r in plan.items[:j])
        src = {"kind": plan.kind(j), "ids": plan.active_ids(j, na), "tiles": plan.tiles(j, na),
               "cma": plan.corner_min_abs(j), "slot": plan._slice(plan.i32, j, "slot"),
               "list_len": plan._slice(plan.i32, j, "list_len"), "lists": plan._slice(plan.u16, j, "lists"),
               "prims": plan.prims[off * 32:(off + len(rec)) * 32], "counts": plan.counts(j)}
        layout, at = {}, 0
        for name, _ in self._PIECES:
            nbytes = src[name].numel() * src[name].element_size()
            layout[name] = (at, nbytes)
            at += (nbytes + 255) // 256 * 256
        dtypes = {"kind": torch.int32, "ids": torch.int32, "tiles": torch.float32, "cma": torch.float32,
                  "slot": torch.int32, "list_len": torch.int32, "lists": torch.int16, "prims": torch.uint8,
                  "counts": torch.int32}
        with torch.cuda.stream(stream):
            self.store = torch.empty(max(at, 256), dtype=torch.uint8, device=plan.device)
        base = self.store.data_ptr()
        segs = []
        for name, _ in self._PIECES:
            o, nbytes = layout[name]
            setattr(self, name, self.store[o:o + nbytes].view(dtypes[name]))
            if name == "counts":
                # [n_active, 0, 0, 0]: the build's count, then zeros (no host round trip)
                segs += [(src[name].data_ptr(), base + o, 4), (0, base + o + 4, nbytes - 4)]
            elif nbytes:
                segs.append((src[name].data_ptr(), base + o, nbytes))
        self.segments = segs
        d = plan.desc_host[j:j + 1].copy()
        d["prims"], d["kind"], d["cma"] = self.prims.data_ptr(), self.kind.data_ptr(), self.cma.data_ptr()
        d["lists"], d["list_len"], d["ids"] = self.lists.data_ptr(), self.list_len.data_ptr(), self.ids.data_ptr()
        d["slot"], d["tiles"], d["counts"] = self.slot.data_ptr(), self.tiles.data_ptr(), self.counts.data_ptr()
        d["tile_cap"] = na
        self.desc_host = d
        self.desc = None
        if copy:
            copy_segments(segs, plan.device, stream)
        if upload:
            with torch.cuda.stream(stream):
                self._set_desc(torch.from_numpy(d.view(np.uint8).copy()).to(plan.device, non_blocking=False))

    def _set_desc(self, desc):
        self.desc = desc
        self.nbytes = self.store.numel() + desc.numel()

    def query_device(self, points, out=None, stream=None):
        """Point SDF at a device float32 (n,3) tensor."""
        torch = _torch()
        if out is None:
            out = torch.empty(points.shape[0], dtype=torch.float32, device=points.device)
        F._check(_lib.lib().fe_query_points(_lib.ptr(self.desc), _lib.ptr(points), points.shape[0], _lib.ptr(out),
                                            _lib.stream_handle(stream)), "point query")
        return out

    def sparse_field(self):
        kind = self.kind.cpu().numpy()
        return F.SparseField(meta=self.meta, kind=kind, active_ids=self.ids.cpu().numpy(),
                             tiles=self.tiles.cpu().numpy(), corner_min_abs=self.cma.cpu().numpy(),
                             stats=F._stats(self.meta, kind, self.n_active))


def copy_segments(segs, device, stream):
    """Device copies [(src address, dst address, bytes)] (src 0: clear) in
    launches of at most 65535 segments, on `stream`."""
    torch = _torch()
    for i in range(0, len(segs), 65535):
        part = np.array(segs[i:i + 65535], SEG)
        with torch.cuda.stream(stream):
            table = torch.from_numpy(part.view(np.uint8).copy()).to(device, non_blocking=False)
        F._check(_lib.lib().fe_copy_segments(_lib.ptr(table), len(part), _lib.stream_handle(stream)),
                 "segment copies")


def resident_fields(plan, stream):
    """ResidentFields of every build of a finished plan: one device-to-host
    read of the active-block counts, one launch copying every field's arrays
    out of the plan's pools, and one upload of the descriptors, all on
    `stream` (so they wait for this plan's work only)."""
    torch = _torch()
    n = len(plan.items)
    with torch.cuda.stream(stream):
        nas = torch.cat([plan.counts(j)[:1] for j in range(n)]).cpu().tolist()
    fields = [ResidentField(plan, j, stream, n_active=nas[j], upload=False, copy=False) for j in range(n)]
    copy_segments([g for f in fields for g in f.segments], plan.device, stream)
    descs = np.concatenate([f.desc_host for f in fields])
    size = descs.dtype.itemsize
    with torch.cuda.stream(stream):
        dev = torch.from_numpy(descs.view(np.uint8).copy()).to(plan.device, non_blocking=False)
    for j, f in enumerate(fields):
        f._set_desc(dev[j * size:(j + 1) * size])
    return fields


def workspace_bytes(meta, recipe):
    """Device bytes of one build's FieldPlan slice (worst-case tile capacity)."""
    SB = F.SUPER[meta.block]
    nsb = -(-meta.n_bx // SB) * -(-meta.n_by // SB) * -(-meta.n_bz // SB)
    cap = min(len(recipe), F.LIST_CAP)
    nb = meta.n_blocks
    return 4 * (3 * nb + nsb + 4) + 4 * (nb + 64 + nb * meta.block ** 3) + 2 * nsb * cap + 32 * len(recipe) + 160


def resident_estimate(meta, recipe):
    """Upper bound of a resident field's bytes before its size is known (all blocks active)."""
    return workspace_bytes(meta, recipe)


class _Request:
    __slots__ = ("rid", "meta", "recipe", "t_arrival", "t_admit", "t_done", "field", "error", "event", "ws", "res",
                 "need")

    def __init__(self, rid, meta, recipe, ws, res):
        self.rid, self.meta, self.recipe = rid, meta, recipe
        self.ws, self.res, self.need = ws, res, ws + res
        self.t_arrival = time.perf_counter()
        self.t_admit = self.t_done = None
        self.field = self.error = None
        self.event = threading.Event()


class FieldServer:
    """Continuous-batching field builds with resident results.

    hbm_budget: bytes of device memory the server may hold: workspaces of
    in-flight batches, the workspaces slots keep between batches, and
    resident fields.  slots: batches in flight at once (one stream each);
    each slot keeps its last workspace pools between batches (reused when
    large enough, so steady serving allocates nothing per batch; counted in
    `reserved` while kept).  A batch is charged its resident estimates, its
    recipe upload and the growth of the slot's pools; a kept pool too small
    for the batch is released before its replacement is allocated, so a slot
    never holds both.  max_batch: requests per batch."""

    def __init__(self, hbm_budget=200e9, slots=2, max_batch=8, device=None, start=True):
        torch = _torch()
        self.device = torch.device(device or "cuda:0")
        self.budget = float(hbm_budget)
        self.max_batch = max_batch
        self.slots = slots
        self.streams = []
        self.free_streams = []
        self.pending = collections.deque()
        self.inflight = []
        self.fields = {}
        self.requests = {}
        self.reserved = 0.0  # workspaces in flight and kept + resident fields
        self.slot_pools = {}  # id(stream) -> the FieldPlan pools last used on that stream
        self.retained = {}  # id(stream) -> {pool name: bytes} of those pools (part of reserved)
        self.ids = itertools.count()
        self.lock = threading.Condition()
        self.running = True
        self.completed = []
        self.query_ms = []
        self.worker = threading.Thread(target=self._loop, name="fe-field-server", daemon=True)
        if start:
            self.worker.start()

    # ------------------------------------------------------------- client side
    def submit(self, meta, recipe):
        """Queues a build; returns its request id.  Raises ValueError when one
        request alone exceeds the HBM budget (admission refusal)."""
        recipe = R.validate(recipe)
        ws, res = workspace_bytes(meta, recipe), resident_estimate(meta, recipe)
        need = ws + res
        if need > self.budget:
            raise ValueError(f"request needs {need / 1e9:.2f} GB, above the {self.budget / 1e9:.2f} GB HBM budget")
        with self.lock:
            if not self.running:
                raise RuntimeError("server closed")
            rid = next(self.ids)
            req = _Request(rid, meta, recipe, ws, res)
            self.requests[rid] = req
            self.pending.append(req)
            self.lock.notify_all()
        return rid

    def result(self, rid, timeout=None):
        req = self.requests[rid]
        if not req.event.wait(timeout):
            raise TimeoutError(f"request {rid} not finished")
        if req.error is not None:
            raise RuntimeError(f"request {rid} failed: {req.error}")
        return req.field

    def query(self, rid, points):
        """Point SDF of resident field `rid` at host points (n,3) -> float32 (n)."""
        torch = _torch()
        f = self.result(rid)
        t0 = time.perf_counter()
        pts = torch.from_numpy(np.ascontiguousarray(points, np.float32)).to(self.device)
        out = f.query_device(pts).cpu().numpy()
        self.query_ms.append((time.perf_counter() - t0) * 1e3)
        return out

    def evict(self, rid):
        with self.lock:
            f = self.fields.pop(rid, None)
            if f is not None:
                self.reserved -= f.nbytes
                self.requests[rid].field = None
                self.lock.notify_all()

    def close(self):
        with self.lock:
            self.running = False
            self.lock.notify_all()
        if self.worker.is_alive():
            self.worker.join()
        with self.lock:
            self.slot_pools.clear()
            self.reserved -= sum(sum(k.values()) for k in self.retained.values())
            self.retained.clear()

    def stats(self):
        b = np.array([(r.t_done - r.t_arrival) * 1e3 for r in self.completed]) if self.completed else np.zeros(1)
        q = np.array(self.query_ms) if self.query_ms else np.zeros(1)
        return dict(completed=len(self.completed), build_p50_ms=float(np.percentile(b, 50)),
                    build_p99_ms=float(np.percentile(b, 99)), query_p50_ms=float(np.percentile(q, 50)),
                    query_p99_ms=float(np.percentile(q, 99)), resident_bytes=self.reserved)

    # ------------------------------------------------------------- worker
    @staticmethod
    def _charge(reqs, kept):
        """Bytes a batch of requests adds to `reserved` on a slot keeping the
        pools `kept` ({name: bytes}): each pool's growth (a kept pool at least
        as large is reused, a smaller one is released before its replacement
        is allocated), the recipe upload and the resident estimates."""
        need = F.plan_pool_bytes([(r.meta, r.recipe) for r in reqs])
        names = set(need) | set(kept)
        grow = sum(max(need.get(k, 0), kept.get(k, 0)) - kept.get(k, 0) for k in names)
        return grow + 32 * sum(len(r.recipe) for r in reqs) + sum(r.res for r in reqs)

    def _admit(self, kept=None):
        """Pops the longest FIFO prefix that fits one batch and the budget on
        a slot keeping the pools `kept`; (batch, charge)."""
        kept = kept or {}
        batch, charge = [], 0.0
        while self.pending and len(batch) < self.max_batch:
            req = self.pending[0]
            if batch and req.meta.block != batch[0].meta.block:
                break
            c = self._charge(batch + [req], kept)
            if self.reserved + c > self.budget:
                break
            batch.append(self.pending.popleft())
            charge = c
        return batch, charge

    def _start(self, key, charge):
        """A batch admitted on slot `key`: `reserved` grows by its charge and
        the slot's kept pools move inside the batch's in-flight charge, which
        this returns."""
        kept = self.retained.pop(key, {})
        self.reserved += charge
        return charge + sum(kept.values())

    def _settle(self, key, held):
        """A batch on slot `key` ended: its in-flight charge `held` leaves
        `reserved` and the pools the slot now keeps enter it."""
        kept = {k: t.numel() * t.element_size() for k, t in self.slot_pools.get(key, {}).items()}
        self.reserved += sum(kept.values()) - held
        self.retained[key] = kept

    def _loop(self):
        torch = _torch()
        torch.cuda.set_device(self.device)
        with self.lock:
            self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.slots)]
            self.free_streams = list(self.streams)
        while True:
            with self.lock:
                while self.running and not (self.pending and self.free_streams) and not self.inflight:
                    self.lock.wait(0.05)
                if not self.running and not self.pending and not self.inflight:
                    return
                batch, held = [], 0.0
                if self.free_streams and self.pending:
                    stream = self.free_streams[-1]
                    batch, charge = self._admit(self.retained.get(id(stream)))
                    if batch:
                        self.free_streams.pop()
                        held = self._start(id(stream), charge)
            if batch:
                now = time.perf_counter()
                for r in batch:
                    r.t_admit = now
                try:
                    with torch.cuda.stream(stream):
                        # the slot's workspace of its earlier batches (finished and
                        # harvested before the slot was freed) is reused when large enough
                        plan = F.FieldPlan([(r.meta, r.recipe) for r in batch], device=self.device,
                                           pools=self.slot_pools.get(id(stream)))
                        with self.lock:
                            self.slot_pools[id(stream)] = plan.pools
                        plan.launch(stream)
                        ev = torch.cuda.Event()
                        ev.record(stream)
                    self.inflight.append((plan, batch, held, stream, ev))
                except Exception as e:  # admission-time failure completes the requests with an error
                    self._fail(batch, held, stream, e)
            self._harvest(block=not batch)

    def _fail(self, batch, held, stream, err):
        with self.lock:
            self._settle(id(stream), held)
            self.free_streams.append(stream)
        for r in batch:
            r.error = repr(err)
            r.t_done = time.perf_counter()
            r.event.set()

    def _harvest(self, block):
        torch = _torch()
        if block and not self.inflight:
            with self.lock:  # nothing on the device: wait for arrivals or evictions
                self.lock.wait(0.001)
            return
        done = [x for x in self.inflight if x[4].query()]
        if not done and block and self.inflight:
            self.inflight[0][4].synchronize()
            done = [self.inflight[0]]
        for item in done:
            self.inflight.remove(item)
            plan, batch, held, stream, ev = item
            fields = []
            try:
                fields = resident_fields(plan, stream)
                stream.synchronize()
                err = None
            except Exception as e:
                err = e
            del plan
            with self.lock:
                self._settle(id(stream), held)
                self.free_streams.append(stream)
                now = time.perf_counter()
                for r, fld in itertools.zip_longest(batch, fields):
                    if err is None:
                        r.field = fld
                        self.fields[r.rid] = fld
                        self.reserved += fld.nbytes
                    else:
                        r.error = repr(err)
                    r.t_done = now
                    self.completed.append(r)
                    r.event.set()
                self.lock.notify_all()


def poisson_run(server, items, rate, queries_per_field=100000, seed=5, query=True):
    """Submits `items` [(meta, recipe)] with Poisson inter-arrival times at
    `rate` requests/s, then queries each field with uniform points in its
    bounding box.  Returns (fields/s over the run, server stats)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(items))
    t0 = time.perf_counter()
    rids = []
    next_t = t0
    for (meta, rec), gap in zip(items, gaps):
        next_t += gap
        while time.perf_counter() < next_t:
            time.sleep(min(0.0005, max(0.0, next_t - time.perf_counter())))
        rids.append(server.submit(meta, rec))
    for rid in rids:
        server.result(rid)
    elapsed = time.perf_counter() - t0
    if query:
        for rid, (meta, rec) in zip(rids, items):
            lo = np.array([meta.x0, meta.y0, meta.z0])
            hi = lo + (np.array([meta.nx, meta.ny, meta.nz]) - 1) * meta.pitch
            server.query(rid, rng.uniform(lo, hi, (queries_per_field, 3)))
    return len(items) / elapsed, server.stats(), rids
