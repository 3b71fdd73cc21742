"""Block-sparse field build on the gfx950 library, behind the reference's entry points.

Reference entry points kept by name, arguments and return values:
  GridMeta, SparseField                              FE/faltkarna_v1.py:50-74
  klassificera_och_evaluera_testfalt                 FE/faltkarna_v1.py:258-299 (plain per-call path)
  rekonstruera_tat / rekonstruera_tat_gpu            FE/faltkarna_v1.py:302-350
  klassificera_evaluera_rekonstruera_gpu_resident    FE/faltkarna_v1_g7_fix.py:56-112 (resident path)
  volym_via_gpu_reduktion, determinism_cert,
  konservering_check                                 FE/faltkarna_v1.py:353-415
  klassificera_och_evaluera_fran_tatt_falt           FE/faltkarna_v1_mesh_to_sdf.py:660-692
  reconstruct_solid_from_sparse                      FE/faltkarna_v1_mesh_to_sdf.py:695-710
The *_recept variants take any recipe (recipes.py) instead of the test field.

FieldPlan owns the device buffers of one or many builds that share a block
edge and launches classify -> compaction -> tiles with no host sync; the
batched builder and the server capture it in a HIP graph.
"""
import ctypes as ct
import math
import os
import subprocess
import sys
import time
from dataclasses import dataclass, field as dc_field

import numpy as np

from . import _lib
from . import recipes as R

BIG = 1.0e6
FIXPT_SCALE = 1.0e6
SUPER = {4: 8, 8: 4, 16: 4}  # blocks per superblock edge, by block edge
LIST_CAP = 512

FIELD_DESC = np.dtype([
    ("x0", "<f4"), ("y0", "<f4"), ("z0", "<f4"), ("pitch", "<f4"),
    ("margin", "<f4"), ("slack", "<f4"), ("nprims", "<i4"), ("B", "<i4"),
    ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("SB", "<i4"),
    ("nbx", "<i4"), ("nby", "<i4"), ("nbz", "<i4"), ("cap", "<i4"),
    ("nsbx", "<i4"), ("nsby", "<i4"), ("nsbz", "<i4"), ("corrupt", "<i4"),
    ("prims", "<u8"), ("kind", "<u8"), ("cma", "<u8"), ("lists", "<u8"), ("list_len", "<u8"),
    ("ids", "<u8"), ("slot", "<u8"), ("tiles", "<u8"), ("counts", "<u8"), ("tile_cap", "<i8")])
assert FIELD_DESC.itemsize == 160


@dataclass
class GridMeta:
    """FE/faltkarna_v1.py:50-64: block counts and the Lipschitz margin sqrt(3)*B*h."""
    x0: float
    y0: float
    z0: float
    pitch: float
    nx: int
    ny: int
    nz: int
    block: int = 8

    def __post_init__(self):
        self.n_bx = -(-self.nx // self.block)
        self.n_by = -(-self.ny // self.block)
        self.n_bz = -(-self.nz // self.block)
        self.n_blocks = self.n_bx * self.n_by * self.n_bz
        self.margin_mm = math.sqrt(3.0) * self.block * self.pitch


@dataclass
class SparseField:
    """FE/faltkarna_v1.py:67-74."""
    meta: GridMeta
    kind: np.ndarray
    active_ids: np.ndarray
    tiles: np.ndarray
    corner_min_abs: np.ndarray
    stats: dict = dc_field(default_factory=dict)


def _torch():
    import torch
    return torch


def _check(status, what):
    if status:
        detail = _lib.lib().fe_last_error().decode() if status == 2 else ""
        raise RuntimeError(f"{what} failed with status {status} {detail}")


def _slack(meta, recipe):
    m = max(abs(meta.x0), abs(meta.y0), abs(meta.z0),
            abs(meta.x0 + meta.n_bx * meta.block * meta.pitch), abs(meta.y0 + meta.n_by * meta.block * meta.pitch),
            abs(meta.z0 + meta.n_bz * meta.block * meta.pitch))
    return 1.0e-4 * (1.0 + m + R.extent(recipe))


def _validate_meta(meta):
    if meta.block not in SUPER:
        raise ValueError("Block edge must be 4, 8 or 16")
    if min(meta.nx, meta.ny, meta.nz) < 1 or not (math.isfinite(meta.pitch) and meta.pitch > 0):
        raise ValueError("Positive grid required")
    if meta.n_blocks >= 2 ** 31 or max(meta.n_bx, meta.n_by, meta.n_bz) * meta.block > 2 ** 20:
        raise ValueError("Grid too large")


POOL_ITEM_BYTES = {"i32": 4, "f32": 4, "u16": 2}


def _layout(items, caps=None):
    """Sizes of a plan over items [(GridMeta, recipe)]: superblock counts,
    list capacities, the padded cma lengths and the element count of each
    pooled allocation (int32: kind, ids, slots, list lengths, counters;
    float32: cma then tiles, the tiles 256-byte aligned as the reconstruction
    reads them as float4; uint16: culled lists)."""
    B = items[0][0].block
    SB = SUPER[B]
    caps = caps if caps is not None else [m.n_blocks for m, _ in items]
    sbs = [(-(-m.n_bx // SB), -(-m.n_by // SB), -(-m.n_bz // SB)) for m, _ in items]
    nsb = [a * b * c for a, b, c in sbs]
    lcap = [min(len(r), LIST_CAP) for _, r in items]
    cma_len = [-(-m.n_blocks // 64) * 64 for m, _ in items]
    per = dict(i32=[3 * m.n_blocks + nsb[i] + 4 for i, (m, _) in enumerate(items)],
               f32=[cma_len[i] + caps[i] * B ** 3 for i in range(len(items))],
               u16=[nsb[i] * lcap[i] for i in range(len(items))])
    elems = dict(i32=sum(per["i32"]), f32=sum(per["f32"]), u16=max(1, sum(per["u16"])))
    return dict(sbs=sbs, nsb=nsb, lcap=lcap, cma_len=cma_len, per_item=per, pool_elems=elems)


def plan_pool_bytes(items):
    """{pool name: bytes} of a plan over items at full tile capacity."""
    return {k: v * POOL_ITEM_BYTES[k] for k, v in _layout(items)["pool_elems"].items()}


class FieldPlan:
    """Device buffers and descriptor table for R builds with one block edge.

    items: [(GridMeta, recipe)] (+ optional per-item tile capacity in blocks;
    the default is every block, which can never overflow).  pools: optional
    {"i32", "f32", "u16"} device tensors of an earlier plan whose work is
    finished; a pool at least as large as this plan needs is reused (sliced)
    instead of allocating, a smaller one is removed from the dict (released)
    before its replacement is allocated.  self.pools holds the tensors this
    plan uses."""

    def __init__(self, items, tile_caps=None, device=None, corrupt=False, pools=None):
        torch = _torch()
        if not items:
            raise ValueError("At least one build required")
        self.device = torch.device(device or "cuda:0")
        B = items[0][0].block
        for meta, rec in items:
            _validate_meta(meta)
            if meta.block != B:
                raise ValueError("A plan shares one block edge")
        self.B, self.SB = B, SUPER[B]
        self.items = [(m, R.validate(r)) for m, r in items]
        self.max_prims = max(len(r) for _, r in self.items)
        n = len(items)
        caps = tile_caps if tile_caps is not None else [m.n_blocks for m, _ in items]
        b3 = B ** 3
        L = _layout(items, caps)
        sbs, nsb, lcap, cma_len = L["sbs"], L["nsb"], L["lcap"], L["cma_len"]
        nch = [-(-m.n_blocks // 1024) for m, _ in items]
        prim_counts = [len(r) for _, r in items]
        self.pools = {}

        def pool(name, size, dtype):
            t = (pools or {}).get(name)
            if t is not None and (t.dtype != dtype or t.device != self.device or t.numel() < size):
                del pools[name]
                t = None
            if t is None:
                t = torch.empty(size, dtype=dtype, device=self.device)
            self.pools[name] = t
            return t[:size]

        self.i32 = pool("i32", L["pool_elems"]["i32"], torch.int32)
        self.f32 = pool("f32", L["pool_elems"]["f32"], torch.float32)
        self.u16 = pool("u16", L["pool_elems"]["u16"], torch.int16)
        host_prims = np.concatenate([r for _, r in items])
        self.prims = torch.from_numpy(host_prims.view(np.uint8).copy()).to(self.device)
        desc = np.zeros(n, FIELD_DESC)
        oi = of = ou = op = 0
        bi32, bf32, bu16, bpr = self.i32.data_ptr(), self.f32.data_ptr(), self.u16.data_ptr(), self.prims.data_ptr()
        self.views = []
        for i, (meta, rec) in enumerate(self.items):
            d = desc[i]
            d["x0"], d["y0"], d["z0"], d["pitch"] = meta.x0, meta.y0, meta.z0, meta.pitch
            d["margin"] = np.float32(meta.margin_mm)
            d["slack"] = _slack(meta, rec)
            d["nprims"], d["B"] = len(rec), B
            d["nx"], d["ny"], d["nz"], d["SB"] = meta.nx, meta.ny, meta.nz, self.SB
            d["nbx"], d["nby"], d["nbz"], d["cap"] = meta.n_bx, meta.n_by, meta.n_bz, lcap[i]
            d["nsbx"], d["nsby"], d["nsbz"] = sbs[i]
            d["corrupt"] = 1 if corrupt else 0
            nb = meta.n_blocks
            v = dict(kind=(oi, nb), ids=(oi + nb, nb), slot=(oi + 2 * nb, nb), list_len=(oi + 3 * nb, nsb[i]),
                     counts=(oi + 3 * nb + nsb[i], 4), cma=(of, nb), tiles=(of + cma_len[i], caps[i] * b3),
                     lists=(ou, nsb[i] * lcap[i]))
            d["prims"] = bpr + op * 32
            for k in ("kind", "ids", "slot", "list_len", "counts"):
                d[k] = bi32 + 4 * v[k][0]
            d["cma"], d["tiles"] = bf32 + 4 * v["cma"][0], bf32 + 4 * v["tiles"][0]
            d["lists"] = bu16 + 2 * v["lists"][0]
            d["tile_cap"] = caps[i]
            self.views.append(v)
            oi += L["per_item"]["i32"][i]
            of += L["per_item"]["f32"][i]
            ou += L["per_item"]["u16"][i]
            op += prim_counts[i]
        self.desc_host = desc
        self.desc = torch.from_numpy(desc.view(np.uint8).copy()).to(self.device)
        self.sb_prefix = torch.tensor(np.concatenate([[0], np.cumsum(nsb)]).astype(np.int32), device=self.device)
        self.ch_prefix = torch.tensor(np.concatenate([[0], np.cumsum(nch)]).astype(np.int32), device=self.device)
        self.total_sb, self.total_ch = int(sum(nsb)), int(sum(nch))
        self.chunk_count = torch.empty(self.total_ch, dtype=torch.int32, device=self.device)
        self.tile_waves = 16384  # waves of the persistent tile kernel (measured best for one 512^3 field)

    def __len__(self):
        return len(self.items)

    def launch(self, stream=None):
        """classify -> compaction -> tiles for every item, asynchronous on `stream`."""
        _check(_lib.lib().fe_field_build(_lib.ptr(self.desc), len(self.items), self.B, _lib.ptr(self.sb_prefix),
                                         self.total_sb, _lib.ptr(self.ch_prefix), self.total_ch,
                                         _lib.ptr(self.chunk_count), self.tile_waves, self.max_prims,
                                         _lib.stream_handle(stream)),
               "field build")

    def launch_counted(self, evals, stream=None):
        """launch() that also adds its primitive- and interval-evaluation counts
        to `evals`, a device int64 tensor of 6 (layout in include/fe_hip.h,
        fe_field_build_counted)."""
        _check(_lib.lib().fe_field_build_counted(_lib.ptr(self.desc), len(self.items), self.B,
                                                 _lib.ptr(self.sb_prefix), self.total_sb, _lib.ptr(self.ch_prefix),
                                                 self.total_ch, _lib.ptr(self.chunk_count), self.tile_waves,
                                                 self.max_prims, _lib.ptr(evals), _lib.stream_handle(stream)),
               "counted field build")

    def _slice(self, pool, i, name):
        off, n = self.views[i][name]
        return pool[off:off + n]

    def kind(self, i):
        return self._slice(self.i32, i, "kind")

    def counts(self, i):
        return self._slice(self.i32, i, "counts")

    def n_active(self, i):
        return int(self.counts(i)[0].item())

    def tiles(self, i, n_active=None):
        B = self.B
        na = self.n_active(i) if n_active is None else n_active
        return self._slice(self.f32, i, "tiles")[:na * B ** 3].view(na, B, B, B)

    def active_ids(self, i, n_active=None):
        na = self.n_active(i) if n_active is None else n_active
        return self._slice(self.i32, i, "ids")[:na]

    def corner_min_abs(self, i):
        return self._slice(self.f32, i, "cma")

    def overflowed(self, i):
        return bool(self.counts(i)[1].item())

    def sparse_field(self, i):
        meta, _ = self.items[i]
        na = self.n_active(i)
        if na > self.views[i]["tiles"][1] // self.B ** 3:
            raise RuntimeError("Tile capacity exceeded")
        kind = self.kind(i).cpu().numpy()
        sf = SparseField(meta=meta, kind=kind, active_ids=self.active_ids(i, na).cpu().numpy(),
                         tiles=self.tiles(i, na).cpu().numpy(), corner_min_abs=self.corner_min_abs(i).cpu().numpy())
        sf.stats = _stats(meta, kind, na)
        return sf

    def reconstruct(self, i, dense=None, occupancy=None, stream=None):
        """Cropped dense fp32 field (and uint8 occupancy) of item i, on the device."""
        torch = _torch()
        meta, _ = self.items[i]
        shape = (meta.nx, meta.ny, meta.nz)
        if dense is None:
            dense = torch.empty(shape, dtype=torch.float32, device=self.device)
        occ_ptr = _lib.ptr(occupancy) if occupancy is not None else None
        _check(_lib.lib().fe_field_reconstruct(_lib.ptr(self.desc), i, dense.numel(), _lib.ptr(dense), occ_ptr,
                                               _lib.stream_handle(stream)), "reconstruct")
        return dense

    def checksum_i64(self, i, scale=FIXPT_SCALE, stream=None):
        torch = _torch()
        out = torch.zeros(1, dtype=torch.int64, device=self.device)
        _check(_lib.lib().fe_field_checksum(_lib.ptr(self.desc), i, float(scale), _lib.ptr(out),
                                            _lib.stream_handle(stream)), "checksum")
        return int(out.item())


def _stats(meta, kind, n_active):
    b3 = meta.block ** 3
    return dict(n_blocks=int(meta.n_blocks), n_active=int(n_active), n_interior=int(np.sum(kind == -1)),
                n_exterior=int(np.sum(kind == 1)), active_voxels=int(n_active * b3),
                dense_voxels=int(meta.nx * meta.ny * meta.nz), margin_mm=meta.margin_mm)


# ---------------------------------------------------------------- plain per-call path
def klassificera_och_evaluera_recept(meta, recipe, corrupt_index=False, device="cuda:0"):
    """Plain per-call build (the structure of FE/faltkarna_v1.py:258-299): classify,
    sync, download kinds, host compaction (np.where), upload ids, evaluate
    tiles, download.  Allocations are per call."""
    torch = _torch()
    plan = FieldPlan([(meta, recipe)], tile_caps=[0], device=device, corrupt=corrupt_index)
    L = _lib.lib()
    st = _lib.stream_handle(None)
    _check(L.fe_field_classify(_lib.ptr(plan.desc), 1, plan.B, _lib.ptr(plan.sb_prefix), plan.total_sb,
                               plan.max_prims, st),
           "classify")
    torch.cuda.synchronize(plan.device)
    kind = plan.kind(0).cpu().numpy()
    cma = plan.corner_min_abs(0).cpu().numpy()
    active_ids = np.where(kind == 0)[0].astype(np.int32)
    n_active = len(active_ids)
    B = meta.block
    if n_active:
        ids_d = torch.from_numpy(active_ids).to(plan.device)
        tiles_d = torch.empty(n_active * B ** 3, dtype=torch.float32, device=plan.device)
        d = plan.desc_host.copy()
        d[0]["tiles"] = tiles_d.data_ptr()
        d[0]["tile_cap"] = n_active
        _check(L.fe_field_tiles_ids(d.ctypes.data, _lib.ptr(ids_d), n_active, st), "tiles")
        torch.cuda.synchronize(plan.device)
        tiles = tiles_d.cpu().numpy().reshape(n_active, B, B, B)
    else:
        tiles = np.zeros((0, B, B, B), np.float32)
    return SparseField(meta=meta, kind=kind, active_ids=active_ids, tiles=tiles, corner_min_abs=cma,
                       stats=_stats(meta, kind, n_active))


def klassificera_och_evaluera_testfalt(meta, hx, hy, hz, r, corrupt_index=False, device="cuda:0"):
    """FE/faltkarna_v1.py:258-299: the box-minus-cylinder test field into a SparseField."""
    return klassificera_och_evaluera_recept(meta, R.testfalt(hx, hy, hz, r), corrupt_index, device)


# ---------------------------------------------------------------- resident path
def klassificera_evaluera_rekonstruera_recept(meta, recipe, corrupt_index=False, device="cuda:0"):
    """Device-resident build and reconstruction; one host transfer (the dense field).
    Returns (dense, timings, stats) like FE/faltkarna_v1_g7_fix.py:56-112."""
    torch = _torch()
    t = {}
    t0 = time.perf_counter()
    plan = FieldPlan([(meta, recipe)], device=device, corrupt=corrupt_index)
    t["t_setup_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    plan.launch()
    dense_d = plan.reconstruct(0)
    torch.cuda.synchronize(plan.device)
    t["t_build_reconstruct_s"] = time.perf_counter() - t1
    t2 = time.perf_counter()
    dense = dense_d.cpu().numpy()
    t["t_final_download_s"] = time.perf_counter() - t2
    t["t_total_s"] = sum(t.values())
    na = plan.n_active(0)
    stats = dict(n_blocks=int(meta.n_blocks), n_active=na, active_voxels=na * meta.block ** 3,
                 dense_voxels=int(meta.nx * meta.ny * meta.nz))
    return dense, t, stats


def klassificera_evaluera_rekonstruera_gpu_resident(meta, hx, hy, hz, r, corrupt_index=False, device="cuda:0"):
    return klassificera_evaluera_rekonstruera_recept(meta, R.testfalt(hx, hy, hz, r), corrupt_index, device)


def _upload_sparse(sf, device):
    """A FieldPlan-shaped descriptor around a host SparseField (for reconstruction)."""
    torch = _torch()
    meta = sf.meta
    plan = FieldPlan([(meta, R.testfalt(1, 1, 1, 1))], tile_caps=[len(sf.active_ids)], device=device)
    kind = plan.kind(0)
    kind.copy_(torch.from_numpy(np.ascontiguousarray(sf.kind, np.int32)))
    slot = np.full(meta.n_blocks, -1, np.int32)
    slot[sf.active_ids] = np.arange(len(sf.active_ids), dtype=np.int32)
    plan._slice(plan.i32, 0, "slot").copy_(torch.from_numpy(slot))
    if len(sf.active_ids):
        plan._slice(plan.f32, 0, "tiles").copy_(torch.from_numpy(np.ascontiguousarray(sf.tiles, np.float32).ravel()))
    c = plan.counts(0)
    c.copy_(torch.tensor([len(sf.active_ids), 0, 0, 0], dtype=torch.int32))
    return plan


def rekonstruera_tat_gpu(sf, device="cuda:0"):
    """FE/faltkarna_v1.py:328-350: dense (nx,ny,nz) fp32, background -BIG/+BIG."""
    plan = _upload_sparse(sf, device)
    return plan.reconstruct(0).cpu().numpy()


rekonstruera_tat = rekonstruera_tat_gpu


def reconstruct_solid_from_sparse(sf, nx, ny, nz, block):
    """FE/faltkarna_v1_mesh_to_sdf.py:695-710: bool occupancy (interior blocks, tiles < 0)."""
    torch = _torch()
    if isinstance(sf, dict):
        meta = GridMeta(0.0, 0.0, 0.0, 1.0, nx, ny, nz, block)
        sf = SparseField(meta=meta, kind=sf["kind"], active_ids=sf["active_ids"], tiles=sf["tiles"],
                         corner_min_abs=np.zeros(len(sf["kind"]), np.float32))
    plan = _upload_sparse(sf, "cuda:0")
    occ = torch.empty((nx, ny, nz), dtype=torch.uint8, device=plan.device)
    plan.reconstruct(0, occupancy=occ)
    return occ.cpu().numpy().astype(bool)


def checksum_i64(sf, scale=FIXPT_SCALE, device="cuda:0"):
    """Deterministic int64 fixed-point sum over tile values < 0 (FE/faltkarna_v1.py:249-255)."""
    plan = _upload_sparse(sf, device)
    return plan.checksum_i64(0, scale)


def konservering_check(volym_matt_mm3, volym_facit_mm3, tol_frac):
    diff = abs(volym_matt_mm3 - volym_facit_mm3) / max(abs(volym_facit_mm3), 1e-9)
    return dict(volym_matt_mm3=volym_matt_mm3, volym_facit_mm3=volym_facit_mm3, diff_frac=diff, tol_frac=tol_frac,
                GRON=bool(diff <= tol_frac))


def determinism_cert(varden):
    return dict(n_korningar=len(varden), varden=varden, bit_identiska=bool(all(v == varden[0] for v in varden)))


def reduktion_checksum_single_process(tiles_npy_path, device="cuda:0"):
    """FE/faltkarna_v1.py:368-386: one run of both checksums over the values < 0
    of a saved tile array: (float32 atomic sum, int64 fixed-point sum / scale)."""
    torch = _torch()
    vals = np.load(tiles_npy_path, allow_pickle=False).reshape(-1).astype(np.float32)
    dev = torch.device(device)
    v = torch.from_numpy(vals).to(dev)
    out32 = torch.zeros(1, dtype=torch.float32, device=dev)
    out64 = torch.zeros(1, dtype=torch.int64, device=dev)
    _check(_lib.lib().fe_checksum_pair(_lib.ptr(v), len(vals), float(FIXPT_SCALE), _lib.ptr(out32), _lib.ptr(out64),
                                       _lib.stream_handle(None)), "checksum pair")
    return float(out32.item()), float(out64.item()) / FIXPT_SCALE


def volym_via_gpu_reduktion(sf, n_korningar=6, tmp_dir="/tmp"):
    """FE/faltkarna_v1.py:389-415: both checksums of the tiles in n separate
    processes; the int64 fixed-point totals must be bit-identical, the float32
    atomic totals are the order-dependent negative control."""
    path = os.path.join(tmp_dir, f"fe_amd_tiles_{os.getpid()}.npy")
    np.save(path, np.ascontiguousarray(sf.tiles, np.float32))
    pkg = os.path.dirname(os.path.abspath(__file__))
    code = ("import sys,importlib.util\n"
            f"spec=importlib.util.spec_from_file_location('fe_amd',{os.path.join(pkg, '__init__.py')!r},"
            f"submodule_search_locations=[{pkg!r}])\n"
            "m=importlib.util.module_from_spec(spec);sys.modules['fe_amd']=m;spec.loader.exec_module(m)\n"
            f"r32,r64=m.field.reduktion_checksum_single_process({path!r})\n"
            "print(repr(r32),repr(r64))\n")
    f32_totals, i64_totals = [], []
    try:
        for _ in range(n_korningar):
            r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
            if r.returncode != 0:
                raise RuntimeError(f"subprocess reduction failed: {r.stderr[-2000:]}")
            parts = r.stdout.strip().splitlines()[-1].split()
            f32_totals.append(float(parts[0]))
            i64_totals.append(float(parts[1]))
    finally:
        os.unlink(path)
    return dict(f32_unsafe=determinism_cert(f32_totals), i64_fixpunkt_safe=determinism_cert(i64_totals),
                n_active_voxels=int(sf.tiles.size), n_processer=n_korningar)


def checksum_tiles(tiles, scale=FIXPT_SCALE):
    """int64 checksum of an arbitrary tile array, on the device."""
    tiles = np.ascontiguousarray(tiles, np.float32)
    b = tiles.shape[-1] if tiles.ndim == 4 else 8
    n = tiles.size // b ** 3
    meta = GridMeta(0.0, 0.0, 0.0, 1.0, b, b, b, b)
    sf = SparseField(meta=meta, kind=np.zeros(1, np.int32), active_ids=np.zeros(0, np.int32),
                     tiles=tiles.reshape(n, b, b, b), corner_min_abs=np.zeros(1, np.float32))
    plan = FieldPlan([(meta, R.testfalt(1, 1, 1, 1))], tile_caps=[n])
    _torch()
    plan._slice(plan.f32, 0, "tiles").copy_(_torch().from_numpy(tiles.ravel()))
    plan.counts(0).copy_(_torch().tensor([n, 0, 0, 0], dtype=_torch().int32))
    return plan.checksum_i64(0, scale)


# ---------------------------------------------------------------- dense-source classify
def klassificera_och_evaluera_fran_tatt_falt(wp, dense, pitch, block, device="cuda:0"):
    """FE/faltkarna_v1_mesh_to_sdf.py:660-692 (the `wp` argument is accepted and ignored)."""
    torch = _torch()
    dense = np.ascontiguousarray(dense, np.float32)
    nx, ny, nz = dense.shape
    n_bx, n_by, n_bz = -(-nx // block), -(-ny // block), -(-nz // block)
    n_blocks = n_bx * n_by * n_bz
    margin = math.sqrt(3.0) * block * pitch
    dev = torch.device(device)
    dense_d = torch.from_numpy(dense).to(dev)
    kind_d = torch.empty(n_blocks, dtype=torch.int32, device=dev)
    L = _lib.lib()
    st = _lib.stream_handle(None)
    _check(L.fe_dense_classify(_lib.ptr(dense_d), nx, ny, nz, block, float(np.float32(margin)), _lib.ptr(kind_d), st),
           "dense classify")
    kind = kind_d.cpu().numpy()
    active_ids = np.where(kind == 0)[0].astype(np.int32)
    n_active = len(active_ids)
    if n_active:
        ids_d = torch.from_numpy(active_ids).to(dev)
        tiles_d = torch.empty(n_active * block ** 3, dtype=torch.float32, device=dev)
        _check(L.fe_dense_tiles(_lib.ptr(dense_d), nx, ny, nz, block, _lib.ptr(ids_d), n_active, _lib.ptr(tiles_d), st),
               "dense tiles")
        tiles = tiles_d.cpu().numpy().reshape(n_active, block, block, block)
    else:
        tiles = np.zeros((0, block, block, block), np.float32)
    return dict(kind=kind, active_ids=active_ids, tiles=tiles, n_bx=n_bx, n_by=n_by, n_bz=n_bz,
                n_blocks=int(n_blocks), n_active=int(n_active), margin_mm=margin)
