"""Mesh stage on the CPU: the oracle against the reference's captured digests
(reports/cuda_columns.json via tests/golden/mesh_stage.*), the package's host
logic (grid window, method selection, validation) and the C ABI's host-side
refusals.  No GPU.

What pins which oracle function:
* or_columns_scan: the 32-box capacity case and both rotated boxes reproduce
  the reference's sha256 of (columns, depths, signs) exactly, which fixes the
  f64 arithmetic without contraction, the face order and the box tables; the
  capacity case also reproduces the 9 depths that differ from the analytic
  10i+-2 by at most 5.7e-14.
* or_column_mask: the plate at 2, 1 and 0.5 mm (the reference's own hits) and
  the rotated boxes reproduce the solid and surface digests; shifting the
  plate's sample z by +-1e-13 changes the 2 mm solid (samples sit on the
  plate's faces), so the key arithmetic is pinned, not just the shape.
* or_gwn: the open and reversed open boxes reproduce the solid digests.
* or_edt_pair: all seven cases reproduce the distance digests.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle as O
from mesh_cases import (capacity_case, column_xy, digest, fixtures, generated_cases, layer_z, oracle_stage,
                        plate_columns, reference_case, window)


@pytest.fixture(scope="module")
def fx():
    return fixtures()


def test_capacity_case_digests(fx):
    meta, arrays = fx
    corners, xy = capacity_case(meta, arrays)
    counts, depths, signs = O.columns_scan(corners, xy, 256)
    c, z, s = O.sorted_hits(counts, depths, signs)
    cap = meta["capacity"]
    assert [digest(c), digest(z), digest(s)] == cap["digests"]
    assert len(c) == cap["hits"] == 64
    analytic = np.array([zz for i in range(32) for zz in (10 * i - 2, 10 * i + 2)], np.float64)
    assert int(np.count_nonzero(z != analytic)) == cap["analytic_depth_differences"]
    assert float(np.max(np.abs(z - analytic))) == cap["analytic_max_delta"]
    assert np.array_equal(s, np.tile([-1, 1], 32))


def test_rotated_hits_and_fields(fx):
    meta, arrays = fx
    for name in ("rotated_0.0", "rotated_1000000.0"):
        v, f, pitch, lo, closed = reference_case(meta, arrays, name)
        want = meta["cases"][name]["digests"]
        st = oracle_stage(v, f, pitch, lo, closed)
        assert list(st["shape"]) == meta["cases"][name]["shape"]
        assert digest(st["gmin"]) == want["gmin"]
        hits = O.sorted_hits(st["counts"], st["depths"], st["signs"])
        assert [digest(h) for h in hits] == [want["columns"], want["depths"], want["signs"]], name
        assert digest(st["solid"]) == want["solid"] and digest(st["surface"]) == want["surface"], name
        assert digest(st["distance"]) == want["distance"], name


def test_plate_mask_from_reference_hits(fx):
    meta, arrays = fx
    for p in ("2.0", "1.0", "0.5"):
        want = meta["cases"][f"plate_{p}"]["digests"]
        counts, depths, signs, shape, zs = plate_columns(meta, arrays, p)
        hits = O.sorted_hits(counts, depths, signs)
        assert [digest(hits[0]), digest(hits[1]), digest(hits[2])] == [want["columns"], want["depths"],
                                                                       want["signs"]]
        solid, surface, flags = O.column_mask(counts, depths, signs, shape, zs)
        assert flags == 3
        assert digest(solid) == want["solid"] and digest(surface) == want["surface"], p
        assert digest(O.edt_pair(solid, float(p))[0]) == want["distance"], p


def test_plate_mask_resolves_1e13(fx):
    """Planted mistakes in the oracle's input: sample z off by +-1e-13, and
    cell-centre samples (z + pitch/2) instead of sample points."""
    meta, arrays = fx
    counts, depths, signs, shape, zs = plate_columns(meta, arrays, "2.0")
    want = meta["cases"]["plate_2.0"]["digests"]["solid"]
    for shifted in (zs + 1e-13, zs - 1e-13, zs + 1.0):
        solid, _, _ = O.column_mask(counts, depths, signs, shape, shifted)
        assert digest(solid) != want


def test_open_boxes_by_winding_number(fx):
    meta, arrays = fx
    for name in ("open", "reversed_open"):
        v, f, pitch, lo, closed = reference_case(meta, arrays, name)
        assert not closed
        want = meta["cases"][name]["digests"]
        st = oracle_stage(v, f, pitch, lo, False)
        assert digest(st["gmin"]) == want["gmin"]
        assert digest(st["solid"]) == want["solid"] and digest(st["surface"]) == want["surface"], name
        assert digest(st["distance"]) == want["distance"], name
        assert np.min(np.abs(np.abs(st["w"]) - .5)) > 1e-3  # no sample near the threshold


def test_gwn_closed_forms():
    """A closed box winds +1 inside and 0 outside; reversing the faces negates w
    (up to the rounding of the reordered denominator product)."""
    meta, arrays = fixtures()
    v, f = arrays["capacity_vertices"], arrays["capacity_faces"]
    rng = np.random.default_rng(4)
    inside = rng.uniform(-1.9, 1.9, (200, 3))
    outside = rng.uniform(2.1, 5.0, (200, 3)) * rng.choice([-1, 1], (200, 3))
    assert np.allclose(O.gwn(v[f], inside), 1.0, atol=1e-12)
    assert np.allclose(O.gwn(v[f], outside), 0.0, atol=1e-12)
    assert np.allclose(O.gwn(v[f[:, ::-1]], inside), -O.gwn(v[f], inside), rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        O.gwn(v[f], v[:1])


def test_scan_brute_force_on_a_tetrahedron():
    """Point-in-triangle by exact rational barycentrics on a tiny mesh."""
    from fractions import Fraction as Fr
    v = np.array([[0.0, 0.0, 0.0], [3.0, 0.5, 0.25], [0.5, 2.5, 0.5], [1.0, 1.0, 2.0]])
    f = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    rng = np.random.default_rng(9)
    xy = rng.uniform(-0.2, 3.2, (400, 2))
    counts, depths, signs = O.columns_scan(v[f], xy, 8)
    for q in range(len(xy)):
        hits = []
        for t in v[f]:
            (ax, ay, az), (bx, by, bz), (cx, cy, cz) = [[Fr(x) for x in p] for p in t]
            qx, qy = Fr(xy[q, 0]), Fr(xy[q, 1])
            d = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
            l0 = (bx - qx) * (cy - qy) - (by - qy) * (cx - qx)
            l1 = (cx - qx) * (ay - qy) - (cy - qy) * (ax - qx)
            l2 = (ax - qx) * (by - qy) - (ay - qy) * (bx - qx)
            if d != 0 and all(l * d >= 0 for l in (l0, l1, l2)):
                hits.append(float((l0 * az + l1 * bz + l2 * cz) / d))
        assert counts[q] == len(hits)
        assert np.allclose(np.sort(depths[q, :counts[q]]), np.sort(hits), rtol=0, atol=1e-14)
        if counts[q] == 2:  # a closed mesh: one entry and one exit
            assert sorted(signs[q, :2]) == [-1, 1]


def test_package_grid_window_and_method(fe, fx):
    meta, arrays = fx
    for name in ("rotated_0.0", "rotated_1000000.0", "open", "reversed_open"):
        v, f, pitch, lo, closed = reference_case(meta, arrays, name)
        gmin, shape = fe.mesh.grid_window(v, pitch, lo)
        assert digest(gmin) == meta["cases"][name]["digests"]["gmin"]
        assert list(shape) == meta["cases"][name]["shape"]
        watertight, winding = fe.mesh.mesh_topology(f)
        method = "cuda_columns" if watertight and winding else "ordered_gwn"
        assert method == meta["cases"][name]["method"], name
    for name, v, f, pitch, lo in generated_cases():
        assert fe.mesh.mesh_topology(f) == (True, True), name
        assert fe.mesh.grid_window(v, pitch, lo)[1] == window(v, pitch, lo)[1]
    # a flipped face breaks winding consistency; a dropped face breaks closure
    v, f = arrays["capacity_vertices"], arrays["capacity_faces"].copy()
    g = f.copy()
    g[3] = g[3, ::-1]
    assert fe.mesh.mesh_topology(g) == (True, False)
    assert fe.mesh.mesh_topology(f[1:])[0] is False


def test_global_shape_clamp(fe):
    v = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [0.0, 5.0, 0.0], [0.0, 0.0, 5.0]])
    gmin, shape = fe.mesh.grid_window(v, 1.0, np.zeros(3), global_shape=(4, 20, 20))
    assert gmin.tolist() == [0, 0, 0] and shape == (4, 8, 8)


def test_prepared_field_validation(fe):
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    f = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    bad = [
        (dict(vertices=np.full((4, 3), np.nan)), 'finite vertices required'),
        (dict(faces=f.astype(np.float64)), 'integral faces required'),
        (dict(faces=f + 1), 'face index outside vertices'),
        (dict(lo=np.zeros(2)), 'finite three-coordinate origin required'),
        (dict(pitch=True), 'bounded real pitch required'),
        (dict(pitch=1e-13), 'bounded real pitch required'),
        (dict(margin_vox_pad=True), 'integer padding required'),
        (dict(margin_vox_pad=513), 'padding outside contract'),
        (dict(global_shape=(4, 4)), 'three integer global dimensions required'),
        (dict(global_shape=(0, 4, 4)), 'global grid outside index contract'),
        (dict(lo=np.array([-1e13, 0, 0])), 'grid coordinates exceed bounded index contract'),
        (dict(faces=np.array([[0, 1, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])), 'degenerate triangle'),
        (dict(pitch=1e-4), 'local grid outside native EDT capacity'),
    ]
    for change, message in bad:
        args = dict(vertices=v, faces=f, pitch=0.1, lo=np.zeros(3))
        args.update(change)
        with pytest.raises(ValueError, match=message):
            fe.PreparedMeshField(args.pop("vertices"), args.pop("faces"), args.pop("pitch"), args.pop("lo"), **args)


def test_abi_host_refusals(fe):
    """Every mesh entry point refuses invalid input on the host (status 1)."""
    lib = fe._lib.lib()
    P = ct.c_void_p
    tri = np.array([[0.0, 0, 0, 1, 0, 0, 0, 1, 0]])
    a = np.zeros(3)
    h = P()
    d = lambda x: x.ctypes.data_as(P)  # noqa: E731
    assert lib.columns_create(b"optix", d(tri), 1, 4, 4, d(a), ct.byref(h)) == 1
    assert lib.columns_create(b"cuda-scan-v1", d(tri), 0, 4, 4, d(a), ct.byref(h)) == 1
    assert lib.columns_create(b"cuda-scan-v1", d(tri), 1, 0, 4, d(a), ct.byref(h)) == 1
    assert lib.columns_create(b"cuda-scan-v1", d(tri), 1, 4, 257, d(a), ct.byref(h)) == 1
    nan = tri.copy()
    nan[0, 4] = np.nan
    assert lib.columns_create(b"cuda-scan-v1", d(nan), 1, 4, 4, d(a), ct.byref(h)) == 1
    assert lib.columns_query(None, d(tri), 1, None, None, None) == 1
    assert lib.columns_destroy(ct.byref(P())) == 0
    keys = np.array([0.5])
    signs = np.array([1], np.int32)
    offs = np.array([0, 1, 2], np.int32)
    fr = np.array([0.5, 0.5])
    out = np.empty(4, np.uint8)
    assert lib.column_mask_surface(d(keys), d(signs), 1, d(offs), d(fr), 2, 1, 2, d(out), d(out)) == 1
    offs = np.array([0, 1, 1], np.int32)
    assert lib.column_mask_surface(d(keys), d(np.array([2], np.int32)), 1, d(offs), d(fr), 2, 1, 2, d(out),
                                   d(out)) == 1
    assert lib.column_mask_surface(d(keys), d(signs), 1, d(offs), d(np.array([0.5, 1.0])), 2, 1, 2, d(out),
                                   d(out)) == 1
    xy = np.zeros((4, 2))
    zs = np.zeros(2)
    assert lib.fused_create(b"", d(tri), 1, d(a), d(xy), 513, 1, 1, 1.0, d(zs), ct.byref(h)) == 1
    assert lib.fused_create(b"", d(tri), 1, d(a), d(xy), 2, 2, 2, 0.0, d(zs), ct.byref(h)) == 1
    assert lib.fe_mesh_create(d(tri), 1, d(a), d(xy), 2, 2, 2, 1.0, d(zs), 0, ct.byref(h)) == 1
    assert lib.fe_mesh_create(d(tri), 1, d(a), d(xy), 4097, 1, 1, 1.0, d(zs), 64, ct.byref(h)) == 1
    pts = np.zeros((8, 3))
    pts[3, 1] = np.inf
    assert lib.fe_gwn_create(d(tri), 1, d(pts), 2, 2, 2, 1.0, ct.byref(h)) == 1
    assert lib.fused_query(None, None, None, None) == 1
    assert lib.fused_destroy(ct.byref(P())) == 0
    assert h.value is None


def test_result_region_layout(fe):
    """FE/mesh_field_shared_service_v1.py:16-19: solid, surface, padding to 64
    bytes, float32 distance."""
    L = fe.mesh_service
    assert L.layout((4, 3, 2)) == (24, 64, 64 + 96)
    assert L.layout((8, 8, 8)) == (512, 1024, 1024 + 2048)
    assert L.layout((1, 1, 1)) == (1, 64, 68)
    assert L.layout((5, 7, 9)) == (315, 640, 640 + 1260)
    region = bytearray(L.layout((5, 7, 9))[2])
    solid, surface, distance = L.result_views(region, (5, 7, 9))
    solid[...] = True
    surface[1, 2, 3] = True
    distance[...] = np.arange(315, dtype=np.float32).reshape(5, 7, 9)
    raw = np.frombuffer(bytes(region), np.uint8)
    assert raw[:315].all() and raw[315:630].sum() == 1 and raw[315 + (1 * 7 + 2) * 9 + 3] == 1
    assert not raw[630:640].any()
    assert np.array_equal(np.frombuffer(bytes(region[640:]), np.float32), np.arange(315, dtype=np.float32))
    ro = L.result_views(bytes(region), (5, 7, 9))
    assert not ro[2].flags.writeable
    with pytest.raises(ValueError):
        L.result_views(bytearray(100), (5, 7, 9))
