"""Per-block pitch: a two-level block-sparse field whose active bricks each
carry their own resolution (FE/faltkarna_v1_multires.py).

The brick lattice is the block lattice of a coarse field (bricks of B^3 coarse
voxels, world side L = B*h).  An active brick at level l in {0, 1, 2} stores
(B*2^l + 1)^3 samples at pitch h/2^l over its closed cube; inactive bricks
store nothing and answer +-margin.  Drop-in entry points, by name and result:

  MultiresField                         FE/faltkarna_v1_multires.py:84-106
  avstand(F, punkter)                   :405-441
  gradient(F, punkter)                  :444-460
  kontinuitetscheck(F, eps_frac, ...)   :462-520
  aktiva_voxlar(F), lagrade_sampel(F)   :366-380

The reference builds from a mesh and picks each brick's level from the local
radius of curvature (trimesh); here the field is a recipe (the fp32 fold of
field.py) and the level map comes from the caller (`bygg_multires_recept`'s
`niva`) or from the recipe's own primitive sizes (`nivaer_fran_recept`, the
same h_l <= r / 10 rule, :160-168).  The map is then dilated by one brick and
2:1 balanced (:171-190).  Level 0 samples sit on the coarse lattice (indices
past the grid clamped to its last sample, as the reference pads its coarse
field), level l on the fine lattice; the overlap band (:211-252) and the
queries interpolate as scipy's map_coordinates(order=1, mode="nearest").
Tiles, bands and queries run on the device (csrc/multires.hip).
"""
import math
from dataclasses import dataclass, field as dc_field

import numpy as np

from . import _lib
from . import field as F_
from . import recipes as R

NIVAER = (0, 1, 2)
MULTIRES_PITCH_FAKTOR = 10.0

MR_BRICK = np.dtype([("bi", "<i4"), ("bj", "<i4"), ("bk", "<i4"), ("level", "<i4"), ("tile", "<i8"),
                     ("item0", "<i8")])
_SLAB_SAMPLES = 1024  # a tile launch's work item: rows of a brick, about this many samples


def slab_rows(m):
    """Rows of one slab of a brick with m samples per axis (csrc/multires.hip,
    mr_slab_rows)."""
    return np.maximum(1, _SLAB_SAMPLES // (m * m))
MR_GRID = np.dtype([("lo", "<f8", (3,)), ("h", "<f8"), ("margin", "<f8"), ("B", "<i4"), ("nb", "<i4", (3,)),
                    ("n", "<i4", (3,)), ("pad", "<i4")])
assert MR_BRICK.itemsize == 32 and MR_GRID.itemsize == 72


def _torch():
    import torch
    return torch


@dataclass
class MultiresField:
    """FE/faltkarna_v1_multires.py:84-106: the brick lattice, kinds and levels
    on the host; tiles on the device (`tile(lin)` downloads one, `tiles` all)."""
    lo: np.ndarray
    h_grov: float
    block: int
    brick_shape: tuple
    kind: np.ndarray
    niva: np.ndarray
    margin_mm: float
    stats: dict = dc_field(default_factory=dict)
    dev: dict = dc_field(default_factory=dict, repr=False)

    @property
    def L(self):
        return self.block * self.h_grov

    def pitch_of(self, lvl):
        return self.h_grov / float(1 << int(lvl))

    def tile(self, lin):
        i = int(self.dev["index"][lin])
        if i < 0:
            raise KeyError(lin)
        b = self.dev["bricks_host"][i]
        m = (self.block << int(b["level"])) + 1
        t = self.dev["tiles"][int(b["tile"]):int(b["tile"]) + m ** 3]
        return t.cpu().numpy().reshape(m, m, m)

    @property
    def tiles(self):
        lins = np.flatnonzero(self.dev["index"] >= 0)
        return {int(l): self.tile(int(l)) for l in lins}


def balansera(niva, aktiv):
    """One brick of level dilation (3x3x3 neighbourhood), then 2:1 balance
    over face neighbours, both among active bricks only (:171-190)."""
    def neighbourhood_max(a, offsets):
        p = np.pad(a, 1, mode="edge")
        out = np.full_like(a, -128)
        nx, ny, nz = a.shape
        for dx, dy, dz in offsets:
            out = np.maximum(out, p[1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny, 1 + dz:1 + dz + nz])
        return out
    cube = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    cross = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    n = np.where(aktiv, niva, -1).astype(np.int8)
    n = np.where(aktiv, np.maximum(n, neighbourhood_max(n, cube)), -1).astype(np.int8)
    for _ in range(len(NIVAER)):
        nxt = np.where(aktiv, np.maximum(n, neighbourhood_max(np.where(aktiv, n, -1).astype(np.int8), cross) - 1),
                       -1).astype(np.int8)
        if np.array_equal(nxt, n):
            break
        n = nxt
    return n


def niva_for_radie(r_mm, h_grov, faktor=MULTIRES_PITCH_FAKTOR):
    """Coarsest level whose pitch is at most r / faktor (the finest when none
    is); (level, clamped)."""
    if r_mm is None or not r_mm > 0.0:
        return 0, False
    for lvl in NIVAER:
        if h_grov / float(1 << lvl) <= r_mm / faktor:
            return lvl, False
    return NIVAER[-1], True


def _feature_radius(recipe):
    """Smallest radius of curvature of each primitive's surface: a sphere's
    radius, a torus's tube radius, a cylinder's radius, a box's edges (0:
    sharp, no refinement by curvature), an infinite cylinder's radius."""
    r = np.asarray(recipe)
    k = r["kind"]
    rad = np.where(k == R.SPHERE, r["a"], 0.0)
    rad = np.where((k >= R.CYL_X) & (k <= R.CYL_Z), r["a"], rad)
    rad = np.where(k == R.TORUS, r["b"], rad)
    rad = np.where(k == R.ZCYL, r["a"], rad)
    return rad.astype(np.float64)


def _prim_boxes(recipe):
    """World boxes (lo, hi) of the primitives' surfaces (the infinite
    cylinder unbounded along z)."""
    r = np.asarray(recipe)
    c = np.stack([r["cx"], r["cy"], r["cz"]], 1).astype(np.float64)
    a, b, cc = (r[k].astype(np.float64) for k in ("a", "b", "c"))
    ext = np.zeros_like(c)
    k = r["kind"]
    ext[k == R.SPHERE] = a[k == R.SPHERE, None]
    ext[k == R.BOX] = np.stack([a, b, cc], 1)[k == R.BOX]
    for ax, kk in ((0, R.CYL_X), (1, R.CYL_Y), (2, R.CYL_Z)):
        sel = k == kk
        e = np.stack([a, a, a], 1)
        e[:, ax] = b
        ext[sel] = e[sel]
    tor = k == R.TORUS
    ext[tor] = np.stack([a + b, a + b, b], 1)[tor]
    zc = k == R.ZCYL
    ext[zc] = np.stack([a, a, np.full_like(a, np.inf)], 1)[zc]
    return c - ext, c + ext


def nivaer_fran_recept(meta, recipe, kind, faktor=MULTIRES_PITCH_FAKTOR):
    """Level per brick from the recipe: a brick's local feature radius is the
    smallest radius of curvature among the curved primitives whose box meets
    the brick; level by niva_for_radie (-1 for inactive bricks)."""
    nb = (meta.n_bx, meta.n_by, meta.n_bz)
    kind = np.asarray(kind).reshape(nb)
    niva = np.full(nb, -1, np.int8)
    act = np.argwhere(kind == 0)
    if not len(act):
        return niva
    h = float(np.float32(meta.pitch))
    L = meta.block * h
    lo0 = np.array([float(np.float32(v)) for v in (meta.x0, meta.y0, meta.z0)])
    blo = lo0 + act * L
    bhi = blo + L
    rad = _feature_radius(recipe)
    plo, phi = _prim_boxes(recipe)
    curved = np.flatnonzero(rad > 0)
    best = np.full(len(act), np.inf)
    for s in range(0, len(curved), 64):
        sel = curved[s:s + 64]
        hit = np.all((blo[:, None, :] <= phi[None, sel, :]) & (bhi[:, None, :] >= plo[None, sel, :]), axis=2)
        r = np.where(hit, rad[None, sel], np.inf).min(axis=1)
        best = np.minimum(best, r)
    for (bi, bj, bk), r in zip(act, best):
        niva[bi, bj, bk] = niva_for_radie(None if not np.isfinite(r) else float(r), h, faktor)[0]
    return niva


def _grid_record(F):
    g = np.zeros((), MR_GRID)
    g["lo"] = F.lo
    g["h"] = F.h_grov
    g["margin"] = F.margin_mm
    g["B"] = F.block
    g["nb"] = F.brick_shape
    g["n"] = F.stats["shape_l"]
    return g


def bygg_multires_recept(meta, recipe, niva=None, overlappsband=True, device="cuda:0"):
    """Builds the per-block-pitch field of a recipe on `meta`'s coarse grid.

    niva: levels per brick (array of the brick lattice's shape, entries of
    inactive bricks ignored) or None for nivaer_fran_recept.  The map is
    dilated and balanced, every active brick's tile is evaluated at its
    pitch, then (overlappsband) the bands of level 1 and then level 2."""
    torch = _torch()
    dev = torch.device(device)
    recipe = R.validate(recipe)
    plan = F_.FieldPlan([(meta, recipe)], tile_caps=[0], device=dev)
    L_ = _lib.lib()
    st = _lib.stream_handle(None)
    F_._check(L_.fe_field_classify(_lib.ptr(plan.desc), 1, plan.B, _lib.ptr(plan.sb_prefix), plan.total_sb,
                                   plan.max_prims, st), "classify")
    nb = (meta.n_bx, meta.n_by, meta.n_bz)
    kind = plan.kind(0).cpu().numpy().reshape(nb).astype(np.int8)
    aktiv = kind == 0
    if niva is None:
        niva = nivaer_fran_recept(meta, recipe, kind)
    niva = np.asarray(niva)
    if niva.shape != nb or not np.all(np.isin(niva[aktiv], NIVAER)):
        raise ValueError(f"level map of shape {nb} with levels {NIVAER} on active bricks required")
    niva = balansera(np.where(aktiv, niva, -1).astype(np.int8), aktiv)
    h = float(np.float32(meta.pitch))
    lo = np.array([float(np.float32(v)) for v in (meta.x0, meta.y0, meta.z0)])
    B = meta.block
    act = np.argwhere(aktiv)
    lin = (act[:, 0] * nb[1] + act[:, 1]) * nb[2] + act[:, 2]
    order = np.argsort(lin)
    act, lin = act[order], lin[order]
    lv = niva[act[:, 0], act[:, 1], act[:, 2]].astype(np.int64)
    m = (B << lv) + 1
    samples = m ** 3
    bricks = np.zeros(len(act), MR_BRICK)
    bricks["bi"], bricks["bj"], bricks["bk"] = act[:, 0], act[:, 1], act[:, 2]
    bricks["level"] = lv
    bricks["tile"] = np.concatenate([[0], np.cumsum(samples)[:-1]]) if len(act) else []
    items = -(-m // slab_rows(m))
    bricks["item0"] = np.concatenate([[0], np.cumsum(items)[:-1]]) if len(act) else []
    index = np.full(int(np.prod(nb)), -1, np.int32)
    index[lin] = np.arange(len(act), dtype=np.int32)
    tiles = torch.empty(max(int(samples.sum()), 1), dtype=torch.float32, device=dev)
    # a field without active bricks still hands the kernels a (never read) table
    d_bricks = torch.from_numpy(bricks.view(np.uint8).copy() if len(act) else np.zeros(32, np.uint8)).to(dev)
    d_slot = torch.from_numpy(index).to(dev)
    d_kind = torch.from_numpy(kind.reshape(-1).astype(np.int32)).to(dev)
    if len(act):
        F_._check(L_.fe_multires_tiles(_lib.ptr(plan.desc), _lib.ptr(d_bricks), len(act), int(items.sum()),
                                       _lib.ptr(tiles), st), "multires tiles")
    F = MultiresField(lo=lo, h_grov=h, block=B, brick_shape=nb, kind=kind, niva=niva,
                      margin_mm=math.sqrt(3.0) * B * h,
                      dev=dict(tiles=tiles, bricks=d_bricks, slot=d_slot, kind=d_kind, index=index,
                               bricks_host=bricks))
    F.stats = dict(shape_l=[meta.nx, meta.ny, meta.nz])
    n_faces = 0
    if overlappsband:
        g = _grid_record(F)
        n_faces = _coarser_faces(F)
        for level in NIVAER[1:]:
            fine = np.flatnonzero(bricks["level"] == level).astype(np.int32)
            if len(fine):
                d_fine = torch.from_numpy(fine).to(dev)
                F_._check(L_.fe_multires_band(g.ctypes.data, _lib.ptr(d_bricks), _lib.ptr(d_slot), _lib.ptr(d_fine),
                                              len(fine), _lib.ptr(tiles), st), "multires band")
    torch.cuda.synchronize(dev)
    per = {int(l): int(np.count_nonzero(niva == l)) for l in NIVAER}
    F.stats.update(n_brickor=int(niva.size), n_aktiva=int(aktiv.sum()), n_per_niva=per,
                   aktiva_voxlar=int(aktiva_voxlar(F)), lagrade_sampel=int(samples.sum()),
                   n_overlappsfaces=int(n_faces), h_grov=h, h_per_niva=[h / (1 << l) for l in NIVAER],
                   origin=[float(x) for x in lo])
    return F


def _coarser_faces(F):
    """Faces between an active brick and an active face neighbour of a lower
    level (the faces the overlap band rewrites)."""
    lv = np.where(F.kind == 0, F.niva, -1).astype(np.int16)
    n = 0
    for ax in range(3):
        a = np.moveaxis(lv, ax, 0)
        lo, hi = a[:-1], a[1:]
        both = (lo >= 0) & (hi >= 0)
        n += int(np.count_nonzero(both & (lo != hi)))
    return n


def aktiva_voxlar(F):
    """Sum over active bricks of (block * 2^level)^3 (:366-374)."""
    return int(sum(int(np.count_nonzero(F.niva == l)) * (F.block * (1 << l)) ** 3 for l in NIVAER))


def lagrade_sampel(F):
    """Samples stored, overlap band included (:377-380)."""
    return int(F.stats["lagrade_sampel"])


def _query(F, pts, with_bricks=False):
    torch = _torch()
    P = np.ascontiguousarray(np.atleast_2d(np.asarray(pts, np.float64)))
    if P.ndim != 2 or P.shape[1] != 3:
        raise ValueError("points of shape (n, 3) required")
    dev = F.dev["tiles"].device
    d_pts = torch.from_numpy(P).to(dev)
    out = torch.empty(len(P), dtype=torch.float64, device=dev)
    bo = torch.empty((len(P), 4), dtype=torch.int32, device=dev) if with_bricks else None
    g = _grid_record(F)
    F_._check(_lib.lib().fe_multires_query(g.ctypes.data, _lib.ptr(F.dev["bricks"]), _lib.ptr(F.dev["slot"]),
                                           _lib.ptr(F.dev["kind"]), _lib.ptr(F.dev["tiles"]), _lib.ptr(d_pts),
                                           len(P), _lib.ptr(out), _lib.ptr(bo) if bo is not None else None,
                                           _lib.stream_handle(None)), "multires query")
    return (out.cpu().numpy(), bo.cpu().numpy() if bo is not None else None)


def avstand(F, punkter):
    """Signed distance (f64) at world points, each at its own brick's pitch."""
    return _query(F, punkter)[0]


def gradient(F, punkter):
    """Central differences at each point's brick pitch, both probes clamped
    into that brick (:444-460); (n, 3).  One query for the bricks, one for
    the six probes of every point."""
    pts = np.atleast_2d(np.asarray(punkter, dtype=np.float64))
    _, bo = _query(F, pts, with_bricks=True)
    b = bo[:, :3].astype(np.float64)
    lvl = bo[:, 3].astype(np.int64)
    h = F.h_grov / np.power(2.0, np.maximum(lvl, 0))
    o = F.lo + b * F.L
    probes = []
    for ax in range(3):
        d = np.zeros((len(pts), 3))
        d[:, ax] = h
        probes += [np.minimum(pts + 0.5 * d, o + F.L), np.maximum(pts - 0.5 * d, o)]
    v = avstand(F, np.concatenate(probes)).reshape(6, len(pts))
    g = np.zeros((len(pts), 3))
    for ax in range(3):
        steg = probes[2 * ax][:, ax] - probes[2 * ax + 1][:, ax]
        g[:, ax] = (v[2 * ax] - v[2 * ax + 1]) / np.maximum(steg, 1e-12)
    return g


def kontinuitetscheck(F, eps_frac=1.0e-3, max_punkter_per_face=None, bara_inre=False):
    """Largest jump of the delivered field across the faces where the pitch
    changes (:462-520): every such face sampled on the finer side's face
    lattice, the query taken eps_frac of the finer pitch on each side; the
    bound is half the coarser pitch.  Every face's points go in one query.

    bara_inre (an extension, default off = the reference's check): only the
    face lattice's interior points, whose two probes stay in the two bricks
    of the face.  A probe from a point on a face's edge can land in a third,
    diagonal brick; when that brick is inactive it answers +-margin, which the
    reference's check then reports as a jump of the field."""
    nb = np.array(F.brick_shape)
    faces, inner, outer = [], [], []
    for b in F.dev["bricks_host"]:
        p = np.array([b["bi"], b["bj"], b["bk"]])
        lvl = int(b["level"])
        n_l = F.block * (1 << lvl)
        h_f = F.pitch_of(lvl)
        b_origin = F.lo + p.astype(np.float64) * F.L
        for ax in range(3):
            for side in (-1, 1):
                q = p.copy()
                q[ax] += side
                if np.any(q < 0) or np.any(q >= nb) or F.kind[tuple(q)] != 0:
                    continue
                lvl_n = int(F.niva[tuple(q)])
                if lvl_n >= lvl:
                    continue
                idx = [np.arange(1 if bara_inre else 0, n_l if bara_inre else n_l + 1, dtype=np.float64)] * 3
                idx[ax] = np.array([0.0 if side < 0 else float(n_l)])
                gi, gj, gk = np.meshgrid(*idx, indexing="ij")
                pts = np.stack([gi.ravel(), gj.ravel(), gk.ravel()], axis=1) * h_f + b_origin
                if max_punkter_per_face is not None and len(pts) > max_punkter_per_face:
                    pts = pts[::int(math.ceil(len(pts) / max_punkter_per_face))]
                e = np.zeros(3)
                e[ax] = eps_frac * h_f
                inner.append(pts - side * e)
                outer.append(pts + side * e)
                faces.append((len(pts), [int(x) for x in p], ax, side, lvl, lvl_n))
    max_hopp, summa, n_pts, varsta, grans = 0.0, 0.0, 0, None, 0.0
    if faces:
        v = avstand(F, np.concatenate(inner + outer))
        half = len(v) // 2
        hopp_all = np.abs(v[:half] - v[half:])
        at = 0
        for cnt, p, ax, side, lvl, lvl_n in faces:
            hopp = hopp_all[at:at + cnt]
            at += cnt
            n_pts += cnt
            summa += float(hopp.sum())
            if float(hopp.max()) > max_hopp:
                max_hopp = float(hopp.max())
                varsta = dict(brick=p, axel=ax, sida=side, niva=lvl, niva_granne=lvl_n)
            grans = max(grans, 0.5 * F.pitch_of(lvl_n))
    n_faces = len(faces)
    return dict(n_pitchbyten_faces=n_faces, n_provpunkter=n_pts, max_hopp_mm=max_hopp,
                medel_hopp_mm=(summa / n_pts if n_pts else 0.0), grans_mm=grans, varsta=varsta,
                GRON=bool(n_faces == 0 or max_hopp <= grans))
