"""GPU parity of the per-block-pitch field (multires.py over csrc/multires.hip)
with the oracle (oracle/multires_ref.py): brick kinds and balanced levels,
every tile bit for bit, with and without the overlap band, and the queries
(avstand, gradient, kontinuitetscheck) at random points, on brick faces and
lattice points, and outside the lattice, bit for bit."""
import math

import numpy as np
import pytest

import multires_ref as MR
import oracle as O

pytestmark = pytest.mark.gpu


def spheres(fe):
    return fe.recipes.make([(0, 0, (12.0, 14.0, 15.0), (7.0,)), (0, 0, (26.0, 22.0, 20.0), (9.0,)),
                            (0, 0, (20.0, 30.0, 28.0), (5.0,))])


def mixed(fe, meta, seed):
    """16 mixed primitives (spheres, boxes, capped cylinders, tori; sizes 0.75
    to 30 voxels) around the middle of meta's grid."""
    rec = fe.recipes.mixed_recipe(np.random.default_rng(seed), 16, 384, spread=0.1)
    o = (meta.x0, meta.y0, meta.z0)
    for a, k in enumerate(("cx", "cy", "cz")):
        rec[k] = (rec[k] - 168.0) * meta.pitch + o[a]
    for k in ("a", "b", "c"):
        rec[k] = rec[k] * meta.pitch
    return rec


def crowded(fe):
    """560 small spheres in one superblock (its culled list over capacity)
    plus 60 mixed primitives over 32^3."""
    from config_cases import crowded_recipe
    return crowded_recipe(fe.recipes, 11, 620, 560, n=32, lo=8.0, hi=24.0)


def slab_levels(nb, axis):
    niva = np.zeros(nb, np.int8)
    sl = [slice(None)] * 3
    sl[axis] = slice(0, 1)
    niva[tuple(sl)] = 2
    return niva


def probe_points(F, rng):
    """Random points in active bricks, brick faces and corners (exact lattice
    coordinates), and points spread over and past the lattice."""
    act = np.argwhere(F.kind == 0)
    lo, L = F.lo, F.L
    a = act[rng.integers(0, len(act), 3000)]
    inside = lo + (a + rng.random((3000, 3))) * L
    b = act[rng.integers(0, len(act), 1500)]
    faces = lo + (b + rng.random((1500, 3))) * L
    ax = rng.integers(0, 3, 1500)
    faces[np.arange(1500), ax] = lo[ax] + (b[np.arange(1500), ax] + rng.integers(0, 2, 1500)) * L
    corners = lo + act[rng.integers(0, len(act), 500)] * L
    ext = np.array(F.kind.shape) * L
    spread = lo - 0.1 * ext + rng.random((2000, 3)) * 1.2 * ext
    fine = lo + (a[:1000] + np.round(rng.random((1000, 3)) * 16) / 16) * L
    return np.concatenate([inside, faces, corners, spread, fine])


def compare(fe, meta, rec, niva, band, what, rng):
    M = fe.multires
    G = M.bygg_multires_recept(meta, rec, niva=niva, overlappsband=band)
    R = MR.Field(meta.x0, meta.y0, meta.z0, meta.pitch, meta.nx, meta.ny, meta.nz, meta.block,
                 np.ascontiguousarray(rec).view(O.PRIM), niva, band=band)
    assert np.array_equal(G.kind, R.kind), what
    assert np.array_equal(G.niva, R.niva), what
    assert np.array_equal(G.lo, R.lo) and G.h_grov == R.h and G.margin_mm == R.margin, what
    got = G.tiles
    assert sorted(got) == sorted(R.tiles), what
    for lin, t in R.tiles.items():
        assert got[lin].tobytes() == t.tobytes(), f"{what}: tile {lin} level {R.level_of(lin)}"
    levels = {int(l): int(np.count_nonzero(R.niva == l)) for l in MR.NIVAER}
    assert G.stats["n_per_niva"] == levels, what
    assert M.lagrade_sampel(G) == sum(t.size for t in R.tiles.values()), what
    pts = probe_points(R, rng)
    assert M.avstand(G, pts).tobytes() == R.avstand(pts).tobytes(), what
    gp = pts[::7]
    assert M.gradient(G, gp).tobytes() == R.gradient(gp).tobytes(), what
    for eps in (1e-3, 1e-6):
        for inner in (False, True):
            c = M.kontinuitetscheck(G, eps_frac=eps, bara_inre=inner)
            faces, npts, mx, bound = R.continuity(eps_frac=eps, interior=inner)
            assert (c["n_pitchbyten_faces"], c["n_provpunkter"], c["max_hopp_mm"], c["grans_mm"]) == \
                (faces, npts, mx, bound), (what, eps, inner)
    return G, R, levels


@pytest.mark.parametrize("n", [40, 38])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("band", [True, False])
def test_spheres_slab_levels(fe, gpu, n, axis, band):
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, n, n, n, block=4)
    nb = (meta.n_bx, meta.n_by, meta.n_bz)
    _, _, levels = compare(fe, meta, spheres(fe), slab_levels(nb, axis), band, f"spheres n{n} ax{axis} band{band}",
                           np.random.default_rng(n * 10 + axis))
    assert all(levels[l] > 0 for l in MR.NIVAER)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mixed_recipe_levels_from_the_recipe(fe, gpu, seed):
    """Levels picked from the primitives' radii, an offset origin and a
    non-unit pitch, block 8."""
    n = 48
    meta = fe.GridMeta(-3.25, 1.5, -0.75, 0.75, n, n + 5, n - 3, block=8)
    rec = mixed(fe, meta, seed)
    g = O.grid_record(np.float32(meta.x0), np.float32(meta.y0), np.float32(meta.z0), np.float32(meta.pitch),
                      meta.nx, meta.ny, meta.nz, meta.block, np.float32(meta.margin_mm))
    kind = O.build_sparse(np.ascontiguousarray(rec).view(O.PRIM), g)[0]
    niva = fe.multires.nivaer_fran_recept(meta, rec, kind)
    G, _, levels = compare(fe, meta, rec, niva, True, f"mixed seed {seed}", np.random.default_rng(seed))
    assert sum(levels[l] for l in (1, 2)) > 0
    # the default level map is the same one
    D = fe.multires.bygg_multires_recept(meta, rec)
    assert np.array_equal(D.niva, G.niva)


def test_crowded_superblock_list(fe, gpu):
    """Tiles of a superblock whose culled list is over capacity fold the full
    recipe."""
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 32, 32, 32, block=4)
    rec = crowded(fe)
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    assert (plan._slice(plan.i32, 0, "list_len").cpu().numpy() == -1).any()
    nb = (meta.n_bx, meta.n_by, meta.n_bz)
    compare(fe, meta, rec, slab_levels(nb, 2), True, "crowded", np.random.default_rng(5))


def test_bad_level_map_and_empty_field(fe, gpu):
    M = fe.multires
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 40, 40, 40, block=4)
    nb = (meta.n_bx, meta.n_by, meta.n_bz)
    with pytest.raises(ValueError):
        M.bygg_multires_recept(meta, spheres(fe), niva=np.full(nb, 3, np.int8))
    with pytest.raises(ValueError):
        M.bygg_multires_recept(meta, spheres(fe), niva=np.zeros((2, 2, 2), np.int8))
    # a sphere far outside: no active brick, every query answers +margin
    far = fe.recipes.make([(0, 0, (500.0, 500.0, 500.0), (3.0,))])
    G = M.bygg_multires_recept(meta, far)
    assert not (G.kind == 0).any() and G.tiles == {}
    v = M.avstand(G, np.array([[5.0, 5.0, 5.0], [-1.0, 0.0, 0.0]]))
    assert np.all(v == G.margin_mm)
    c = M.kontinuitetscheck(G)
    assert c["n_pitchbyten_faces"] == 0 and c["GRON"]
