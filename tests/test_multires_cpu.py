"""The per-block-pitch oracle (oracle/multires_ref.py) pinned by closed forms
and the package's host logic against it.  No GPU.

* Trilinear interpolation reproduces an affine field exactly, so with every
  tile holding samples of one affine function the oracle's avstand equals it
  everywhere (to fp32 rounding of the samples) at every level, and after the
  overlap band the field is continuous across every pitch change; on the
  built sphere field the band closes the interior face jumps.
* Balancing: levels only grow, active face neighbours differ by at most one
  level, and every active brick is at least the largest caller level in its
  3x3x3 neighbourhood; the package's balansera equals the oracle's
  (scipy grey_dilation) on random maps.
* A point exactly on the face between an inactive and an active brick is
  answered by the active one; points outside the lattice get +margin.
* The level rule h/2^l <= r/10, clamped at the finest level.
"""
import math

import numpy as np
import pytest

import multires_ref as MR
import oracle as O


def small_field(band=True, fine_axis=0, B=4, n=40):
    """Three spheres in a 40^3 grid; level 2 asked for in the first two brick
    slabs along fine_axis, which balancing turns into slabs of levels 2, 1, 0."""
    rec = np.zeros(3, O.PRIM)
    for i, (c, r) in enumerate((((12.0, 14.0, 15.0), 7.0), ((26.0, 22.0, 20.0), 9.0), ((20.0, 30.0, 28.0), 5.0))):
        rec[i] = (0, 0, *c, r, 0.0, 0.0)
    nb = -(-n // B)
    niva = np.zeros((nb, nb, nb), np.int8)
    sl = [slice(None)] * 3
    sl[fine_axis] = slice(0, 2)
    niva[tuple(sl)] = 2
    return MR.Field(0.0, 0.0, 0.0, 1.0, n, n, n, B, rec, niva, band=band), rec, niva


def affine(p):
    return 0.37 * p[..., 0] - 0.21 * p[..., 1] + 0.13 * p[..., 2] - 3.0


def fill(F, fn):
    """Overwrites every tile with fn at its samples' f64 positions."""
    for lin in F.tiles:
        bi, bj, bk = F.brick_of_lin(lin)
        lvl = F.level_of(lin)
        m = (F.block << lvl) + 1
        o = F.lo + np.array([bi, bj, bk], np.float64) * F.L
        g = np.arange(m) * F.pitch_of(lvl)
        X, Y, Z = np.meshgrid(o[0] + g, o[1] + g, o[2] + g, indexing="ij")
        F.tiles[lin] = fn(np.stack([X, Y, Z], -1)).astype(np.float32)


def inside_points(F, rng, count):
    act = np.argwhere(F.kind == 0)
    pick = act[rng.integers(0, len(act), count)]
    return F.lo + (pick + rng.random((count, 3))) * F.L


def test_affine_field_reproduced_and_continuous():
    F, _, _ = small_field(band=False)
    assert len(set(F.level_of(l) for l in F.tiles)) == 3
    fill(F, affine)
    rng = np.random.default_rng(1)
    pts = inside_points(F, rng, 4000)
    assert np.abs(F.avstand(pts) - affine(pts)).max() < 2e-6
    for lvl in (1, 2):
        for lin in sorted(F.tiles):
            if F.level_of(lin) == lvl:
                F._band(lin)
    # eps of 1e-6 pitch: the affine slope adds under 1e-6 to a jump
    faces, npts, mx, bound = F.continuity(eps_frac=1e-6)
    assert faces > 0 and mx < 5e-6 and bound == 0.5


def interior_face_jumps(F, eps_frac=1e-6):
    """Largest jump over the interior points of every pitch-change face (edge
    points can fall into an inactive diagonal brick, which answers +-margin),
    restated here from the face lattice as a check of the oracle's option."""
    mx = 0.0
    for lin in sorted(F.tiles):
        b = np.array(F.brick_of_lin(lin))
        lvl = F.level_of(lin)
        n_l = F.block << lvl
        h_f = F.pitch_of(lvl)
        for ax in range(3):
            for side in (-1, 1):
                nb = b.copy()
                nb[ax] += side
                if np.any(nb < 0) or np.any(nb >= F.nb) or F.kind[tuple(nb)] != 0 or F.niva[tuple(nb)] >= lvl:
                    continue
                idx = [np.arange(1, n_l, dtype=np.float64)] * 3
                idx[ax] = np.array([0.0 if side < 0 else float(n_l)])
                g = np.stack(np.meshgrid(*idx, indexing="ij"), -1).reshape(-1, 3)
                pts = F.lo + b * F.L + g * h_f
                e = np.zeros(3)
                e[ax] = eps_frac * h_f
                mx = max(mx, float(np.abs(F.avstand(pts - side * e) - F.avstand(pts + side * e)).max()))
    return mx


@pytest.mark.parametrize("fine_axis", [0, 1, 2])
def test_band_closes_jumps_of_the_built_field(fine_axis):
    """Without the band the fine side's own samples disagree with the coarse
    side's interpolant by the interpolation error; with it the interior face
    points agree to fp32 rounding (1e-5 on values below 40)."""
    F0, F1 = small_field(band=False, fine_axis=fine_axis)[0], small_field(band=True, fine_axis=fine_axis)[0]
    before, after = interior_face_jumps(F0), interior_face_jumps(F1)
    assert before > 1e-3 and after < 1e-5, (before, after)
    assert F0.continuity(1e-6, interior=True)[2] == before and F1.continuity(1e-6, interior=True)[2] == after


def test_balance_rules_and_package_equal(fe):
    rng = np.random.default_rng(3)
    for trial in range(20):
        shape = tuple(int(x) for x in rng.integers(3, 12, 3))
        aktiv = rng.random(shape) < 0.6
        niva = rng.choice([0, 0, 0, 1, 2], shape).astype(np.int8)
        got = MR.balance(niva, aktiv)
        assert np.array_equal(got, fe.multires.balansera(np.where(aktiv, niva, -1).astype(np.int8), aktiv)), trial
        assert np.all(got[~aktiv] == -1) and np.all(got[aktiv] >= niva[aktiv])
        for ax in range(3):
            a = np.moveaxis(got, ax, 0)
            act = np.moveaxis(aktiv, ax, 0)
            both = act[1:] & act[:-1]
            assert np.all(np.abs(a[1:].astype(int) - a[:-1])[both] <= 1), trial
        pad = np.pad(np.where(aktiv, niva, -1), 1, constant_values=-1)
        for i, j, k in np.argwhere(aktiv):
            assert got[i, j, k] >= pad[i:i + 3, j:j + 3, k:k + 3].max()


def test_face_rule_and_outside():
    F, _, _ = small_field()
    B, L = F.block, F.L
    # an inactive brick with an active neighbour below it along some axis
    found = 0
    for i, j, k in np.argwhere(F.kind != 0):
        for ax in range(3):
            q = [i, j, k]
            q[ax] -= 1
            if q[ax] < 0 or F.kind[tuple(q)] != 0:
                continue
            p = F.lo + (np.array([i, j, k], np.float64) + 0.3) * L
            p[ax] = F.lo[ax] + (i, j, k)[ax] * L  # exactly on the shared face
            b, outside = F.bricks_of(p[None])
            assert not outside[0] and F.kind[tuple(b[0])] == 0
            found += 1
    assert found > 0
    far = np.array([[-5.0, 1.0, 1.0], [1.0, 1e3, 1.0]])
    assert np.all(F.avstand(far) == F.margin)


def test_level_rule(fe):
    M = fe.multires
    h = 1.0
    assert M.niva_for_radie(None, h) == (0, False) and M.niva_for_radie(0.0, h) == (0, False)
    assert M.niva_for_radie(10.0, h) == (0, False)
    assert M.niva_for_radie(9.99, h) == (1, False) and M.niva_for_radie(5.0, h) == (1, False)
    assert M.niva_for_radie(4.99, h) == (2, False) and M.niva_for_radie(2.5, h) == (2, False)
    assert M.niva_for_radie(1.0, h) == (2, True)


@pytest.mark.parametrize("n", [40, 38])
def test_level0_tiles_are_the_single_pitch_tiles(n):
    """Inside the lattice a level-0 tile's first B^3 samples are the
    single-pitch sparse build's tile of that block; past n-1 (ragged lattice)
    they repeat the last in-lattice sample, the edge padding of the coarse
    field (faltkarna_v1_multires.py:305), where the single-pitch tile goes on
    sampling the recipe."""
    F, rec, _ = small_field(band=False, fine_axis=2, n=n)
    B = F.block
    g = O.grid_record(F.x0, F.y0, F.z0, F.pitch, n, n, n, B, np.float32(math.sqrt(3.0) * B))
    kind, _cma, ids, tiles = O.build_sparse(rec, g)
    assert np.array_equal(kind.reshape(F.nb), F.kind)
    seen = ragged = 0
    for t, lin in enumerate(ids):
        if F.level_of(int(lin)) != 0:
            continue
        b = F.brick_of_lin(int(lin))
        valid = [min(B, n - bb * B) for bb in b]
        mine = F.tiles[int(lin)]
        assert np.array_equal(mine[:valid[0], :valid[1], :valid[2]], tiles[t][:valid[0], :valid[1], :valid[2]]), int(lin)
        clamp = [np.minimum(np.arange(B + 1), min(B + 1, n - bb * B) - 1) for bb in b]
        assert np.array_equal(mine, mine[np.ix_(*clamp)]), int(lin)
        seen += 1
        ragged += min(valid) < B
    assert seen > 10 and (ragged > 0) == (n % B != 0)
