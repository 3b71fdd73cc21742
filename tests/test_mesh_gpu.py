"""Mesh stage on the GPU through the C ABI, against the reference's captured
digests and the oracle:

* columns_create/query (the reference's column scan ABI): the 32-box capacity
  case reproduces the reference's hit digests; overflow, thread and input
  refusals; every hit (count, depth bits, sign, face order) equals the oracle
  on generated meshes and query sets, including queries on triangle edges;
* column_mask_surface: the plate at 2, 1, 0.5 mm from the reference's hits
  reproduces the reference's solid and surface digests;
* fused_create/query (the captured stage): the rotated boxes reproduce the
  reference's solid, surface and distance digests on every replay;
* PreparedMeshField: all four meshed reference cases (gmin, surface, solid,
  distance digests, method), generated meshes bit-exact against the oracle's
  whole stage, a grid beyond the reference's 512 cap, and a 400^3 property
  check against the analytic sphere;
* winding numbers within 1e-12 of the oracle's (atan2 differs in the last
  bits between libraries), status latching.
"""
import ctypes as ct
import threading

import numpy as np
import pytest

import oracle as O
from mesh_cases import (capacity_case, column_xy, digest, fixtures, generated_cases, layer_z, oracle_stage,
                        plate_columns, reference_case, star_sphere, torus, window)

pytestmark = pytest.mark.gpu
P = ct.c_void_p


def d(a):
    return a.ctypes.data_as(P)


@pytest.fixture(scope="module")
def fx():
    return fixtures()


class Columns:
    """ctypes driver of columns_create / columns_query[_compact] / columns_destroy."""

    def __init__(self, lib, corners, capacity, hits=64):
        self.lib, self.hits = lib, hits
        self.h = P()
        c = np.ascontiguousarray(corners, np.float64).reshape(-1, 9)
        st = lib.columns_create(b"cuda-scan-v1", d(c), len(c), capacity, hits, d(np.zeros(3)), ct.byref(self.h))
        assert st == 0, st

    def query(self, xy, compact=False):
        q = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        counts = np.full(len(q), -7, np.int32)
        depths = np.zeros(len(q) * self.hits, np.float64)
        signs = np.zeros(len(q) * self.hits, np.int32)
        if compact:
            width = ct.c_uint32(12345)
            st = self.lib.columns_query_compact(self.h, d(q), len(q), d(counts), d(depths), d(signs),
                                                ct.byref(width))
            w = width.value
        else:
            st = self.lib.columns_query(self.h, d(q), len(q), d(counts), d(depths), d(signs))
            w = self.hits
        return st, counts, depths[:len(q) * w].reshape(len(q), w), signs[:len(q) * w].reshape(len(q), w)

    def close(self):
        return self.lib.columns_destroy(ct.byref(self.h))


def assert_hits_equal(got, want, name):
    st, counts, depths, signs = got
    wc, wd, ws = want
    assert st == 0, (name, st)
    assert np.array_equal(counts, wc), name
    for q in np.flatnonzero(wc):
        n = wc[q]
        assert depths[q, :n].tobytes() == wd[q, :n].tobytes(), (name, q)
        assert np.array_equal(signs[q, :n], ws[q, :n]), (name, q)


def test_columns_capacity_case_and_refusals(fe, gpu, fx):
    meta, arrays = fx
    lib = fe._lib.lib()
    corners, xy = capacity_case(meta, arrays)
    h = Columns(lib, corners, 1)
    st, counts, depths, signs = h.query(xy)
    assert st == 0 and counts[0] == 64
    hits = O.sorted_hits(counts, depths, signs)
    assert [digest(a) for a in hits] == meta["capacity"]["digests"]
    errors = []
    t = threading.Thread(target=lambda: errors.append(h.query(xy)[0]))
    t.start()
    t.join()
    assert errors == [3]
    for bad in (np.empty((0, 2)), np.zeros((2, 2)), np.array([[np.nan, 0.0]]), np.array([[2e12, 0.0]])):
        assert h.query(bad)[0] == 1
    assert h.close() == 0
    # 33 boxes: 66 hits > 64 -> status 4, with the counts still written
    more = np.concatenate([corners, corners[-12:] + [0, 0, 10.0]])
    h = Columns(lib, more, 1)
    st, counts, _, _ = h.query(xy)
    assert st == 4 and counts[0] == 66
    assert h.close() == 0


def query_sets(v, f, pitch, lo):
    """Grid columns of the stage, random points, and points on triangle edges and corners."""
    gmin, shape = window(v, pitch, lo)
    anchor = lo + gmin * pitch
    rng = np.random.default_rng(int(len(f)))
    tri = v[f][:300]
    w = rng.uniform(0, 1, (len(tri), 1))
    on_edges = tri[:, 0, :2] * w + tri[:, 1, :2] * (1 - w)
    return {"grid": column_xy(anchor, shape, pitch),
            "random": rng.uniform(v.min(0)[:2] - 1, v.max(0)[:2] + 1, (1000, 2)),
            "edges": np.concatenate([on_edges, tri[:200, 2, :2]])}


def test_columns_match_oracle(fe, gpu, fx):
    meta, arrays = fx
    lib = fe._lib.lib()
    cases = [(n, v, f, p, lo) for n, v, f, p, lo in generated_cases()]
    for name in ("rotated_0.0", "rotated_1000000.0"):
        v, f, p, lo, _ = reference_case(meta, arrays, name)
        cases.append((name, v, f, p, lo))
    for name, v, f, pitch, lo in cases:
        sets = query_sets(v, f, pitch, lo)
        h = Columns(lib, v[f], max(len(q) for q in sets.values()))
        for label, xy in sets.items():
            print("case", name, label, len(xy), flush=True)
            want = O.columns_scan(v[f], xy, 64)
            assert_hits_equal(h.query(xy), want, (name, label))
            # compact rows are *width wide; slots past a count are unspecified
            compact = h.query(xy, compact=True)
            assert compact[2].shape[1] == int(want[0].max())
            assert_hits_equal(compact, want, (name, label, "compact"))
        assert h.close() == 0


def native_mask_inputs(c, z, s, pitch, origin, shape):
    """The host preparation of FE/column_mask_native_v1.py:27-33."""
    order = np.lexsort((z, c))
    c, z = c[order], z[order].astype(np.float64)
    signs = np.ascontiguousarray(s[order], np.int32)
    counts = np.bincount(c, minlength=shape[0] * shape[1])
    offsets = np.ascontiguousarray(np.r_[0, np.cumsum(counts)], np.int32)
    low = float(z.min())
    span = max(float(z.max()) - low, 1e-12)
    keys = np.ascontiguousarray(c.astype(np.float64) + .25 + .5 * (z - low) / span)
    zs = origin[2] + (np.arange(shape[2], dtype=np.float64) + 0.0) * pitch
    fractions = np.ascontiguousarray(np.clip(.25 + .5 * (zs - low) / span, 0, .999))
    return keys, signs, offsets, fractions


def test_column_mask_surface_plate(fe, gpu, fx):
    meta, arrays = fx
    lib = fe._lib.lib()
    for p in ("2.0", "1.0", "0.5"):
        name = f"plate_{p}"
        c = meta["cases"][name]
        shape = tuple(c["shape"])
        print("case", name, shape, flush=True)
        keys, signs, offsets, fractions = native_mask_inputs(
            arrays[f"{name}_columns"].astype(np.int64), arrays[f"{name}_depths"],
            arrays[f"{name}_signs"].astype(np.int32), c["pitch"], np.array([0.0, 0.0, c["anchor_z"]]), shape)
        solid = np.empty(shape, np.uint8)
        surface = np.empty(shape, np.uint8)
        st = lib.column_mask_surface(d(keys), d(signs), len(keys), d(offsets), d(fractions), *shape, d(solid),
                                     d(surface))
        assert st == 0
        assert digest(solid.view(bool)) == c["digests"]["solid"], name
        assert digest(surface.view(bool)) == c["digests"]["surface"], name


def test_fused_abi_rotated_cases(fe, gpu, fx):
    meta, arrays = fx
    lib = fe._lib.lib()
    for name in ("rotated_0.0", "rotated_1000000.0"):
        v, f, pitch, lo, _ = reference_case(meta, arrays, name)
        print("case", name, flush=True)
        gmin, shape = window(v, pitch, lo)
        anchor = np.ascontiguousarray(lo + gmin * pitch)
        corners = np.ascontiguousarray(v[f])
        xy = column_xy(anchor, shape, pitch)
        zs = np.ascontiguousarray(layer_z(anchor, shape[2], pitch))
        h = P()
        assert lib.fused_create(b"unused.ptx", d(corners), len(corners), d(anchor), d(xy), *shape, pitch, d(zs),
                                ct.byref(h)) == 0
        want = meta["cases"][name]["digests"]
        for leg in range(3):
            solid = np.empty(shape, np.uint8)
            surface = np.empty(shape, np.uint8)
            distance = np.empty(shape, np.float32)
            assert lib.fused_query(h, d(solid), d(surface), d(distance)) == 0
            assert digest(solid.view(bool)) == want["solid"], (name, leg)
            assert digest(surface.view(bool)) == want["surface"], (name, leg)
            assert digest(distance) == want["distance"], (name, leg)
        assert lib.fused_destroy(ct.byref(h)) == 0 and h.value is None


def test_prepared_field_reference_cases(fe, gpu, fx):
    meta, arrays = fx
    for name in ("rotated_0.0", "rotated_1000000.0", "open", "reversed_open"):
        v, f, pitch, lo, _ = reference_case(meta, arrays, name)
        print("case", name, flush=True)
        want = meta["cases"][name]["digests"]
        with fe.PreparedMeshField(v, f, pitch, lo) as field:
            assert field.method == meta["cases"][name]["method"]
            for leg in range(2):
                gmin, shape, surface, solid, distance = field.evaluate()
                assert list(shape) == meta["cases"][name]["shape"]
                got = [digest(x) for x in (gmin, surface, solid, distance)]
                assert got == [want["gmin"], want["surface"], want["solid"], want["distance"]], (name, leg)
            errors = []

            def foreign():
                try:
                    field.evaluate()
                except RuntimeError:
                    errors.append(True)
            t = threading.Thread(target=foreign)
            t.start()
            t.join()
            assert errors == [True]
        with pytest.raises(RuntimeError):
            field.evaluate()
    out = fe.surface_raster_and_flood_cuda(*reference_case(meta, arrays, "open")[:4])
    assert digest(out[3]) == meta["cases"]["open"]["digests"]["solid"]


def test_generated_meshes_match_oracle(fe, gpu):
    for name, v, f, pitch, lo in generated_cases():
        print("case", name, len(f), flush=True)
        want = oracle_stage(v, f, pitch, lo, True)
        with fe.PreparedMeshField(v, f, pitch, lo) as field:
            assert field.method == "cuda_columns"
            gmin, shape, surface, solid, distance = field.evaluate()
        assert np.array_equal(gmin, want["gmin"]) and shape == want["shape"], name
        assert np.array_equal(solid, want["solid"]), name
        assert np.array_equal(surface, want["surface"]), name
        assert distance.tobytes() == want["distance"].tobytes(), name


def test_open_generated_mesh_matches_oracle(fe, gpu):
    """A sphere with a cap of faces removed goes the winding-number way."""
    v, f = star_sphere(14, 20, 6.0, (0.3, -0.2, 0.1), seed=2, wobble=0.1)
    keep = v[f].mean(1)[:, 2] < 4.5
    f = f[keep]
    lo = v.min(0) - 1.3
    want = oracle_stage(v, f, 0.5, lo, False)
    margin = np.abs(np.abs(want["w"]) - .5) > 1e-9
    with fe.PreparedMeshField(v, f, 0.5, lo) as field:
        assert field.method == "ordered_gwn"
        gmin, shape, surface, solid, distance = field.evaluate()
    assert margin.all()
    assert np.array_equal(solid, want["solid"]) and np.array_equal(surface, want["surface"])
    assert distance.tobytes() == want["distance"].tobytes()


def test_grid_beyond_reference_cap(fe, gpu):
    """A long torus whose grid has 600+ columns along x (the reference stops at 512)."""
    v, f = torus(9.0, 2.0, 64, 16, (0.0, 0.0, 0.0), tilt=0.3)
    v = v * np.array([14.0, 1.0, 1.0])
    lo = v.min(0) - 0.5
    want = oracle_stage(v, f, 0.25, lo, True)
    assert want["shape"][0] > 512
    print("shape", want["shape"], flush=True)
    with fe.PreparedMeshField(v, f, 0.25, lo) as field:
        gmin, shape, surface, solid, distance = field.evaluate()
    assert shape == want["shape"]
    assert np.array_equal(solid, want["solid"]) and np.array_equal(surface, want["surface"])
    assert distance.tobytes() == want["distance"].tobytes()


def test_large_sphere_properties(fe, gpu):
    """400^3 voxels: samples inside the inscribed radius are solid, samples
    outside the sphere are not, the signed distance is negative exactly on
    solid voxels, and the surface is the solid's 6-neighbour boundary."""
    R = 99.0
    n_lat, n_lon = 96, 192
    v, f = star_sphere(n_lat, n_lon, R, (0.0, 0.0, 0.0))
    pitch = 0.5
    lo = v.min(0) - 1.0
    with fe.PreparedMeshField(v, f, pitch, lo) as field:
        gmin, shape, surface, solid, distance = field.evaluate()
        anchor = field._anchor
    print("shape", shape, flush=True)
    assert np.prod(shape) > 6.0e7
    ax = [anchor[i] + np.arange(shape[i]) * pitch for i in range(3)]
    r2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
    inner = R * np.cos(np.pi / n_lat) * np.cos(np.pi / n_lon) - pitch
    assert solid[r2 < inner ** 2].all()
    assert not solid[r2 > (R + pitch) ** 2].any()
    assert np.array_equal(distance < 0, solid)
    assert np.array_equal(surface, O.surface_of(solid))


def test_gwn_device_values(fe, gpu):
    import torch
    v, f = star_sphere(10, 16, 3.0, (0.0, 0.0, 0.0), seed=8, wobble=0.2)
    f = f[v[f].mean(1)[:, 0] < 1.5]
    rng = np.random.default_rng(6)
    pts = rng.uniform(-4, 4, (20000, 3))
    want = O.gwn(v[f], pts)
    lib = fe._lib.lib()
    tc = torch.tensor(np.ascontiguousarray(v[f]), dtype=torch.float64, device=gpu)
    tp = torch.tensor(pts, dtype=torch.float64, device=gpu)
    w = torch.empty(len(pts), dtype=torch.float64, device=gpu)
    m = torch.empty(len(pts), dtype=torch.uint8, device=gpu)
    flags = torch.zeros(1, dtype=torch.int32, device=gpu)
    st = lib.fe_gwn_device(fe._lib.ptr(tc), len(f), fe._lib.ptr(tp), len(pts), fe._lib.ptr(w), fe._lib.ptr(m),
                           fe._lib.ptr(flags), fe._lib.stream_handle())
    assert st == 0
    torch.cuda.synchronize()
    got = w.cpu().numpy()
    assert np.max(np.abs(got - want)) < 1e-12
    clear = np.abs(np.abs(want) - .5) > 1e-9
    assert np.array_equal(m.cpu().numpy().astype(bool)[clear], (np.abs(want) > .5)[clear])
    assert int(flags.item()) == 3


def test_status_latching(fe, gpu, fx):
    meta, arrays = fx
    lib = fe._lib.lib()
    # hit overflow: 33 stacked boxes under one column -> 4, then latched 2
    corners, _ = capacity_case(meta, arrays)
    corners = np.ascontiguousarray(np.concatenate([corners, corners[-12:] + [0, 0, 10.0]]))
    shape = (2, 2, 8)
    anchor = np.array([0.0, 0.0, -10.0])
    xy = np.ascontiguousarray(np.array([[.123, .234], [.5, .5], [3.0, 3.0], [-1.0, 1.0]]))
    zs = np.ascontiguousarray(np.linspace(-10, 330, 8))
    h = P()
    assert lib.fused_create(b"", d(corners), len(corners), d(anchor), d(xy), *shape, 1.0, d(zs), ct.byref(h)) == 0
    out = [np.empty(shape, np.uint8), np.empty(shape, np.uint8), np.empty(shape, np.float32)]
    assert lib.fused_query(h, *map(d, out)) == 4
    assert lib.fused_query(h, *map(d, out)) == 2
    assert lib.fused_destroy(ct.byref(h)) == 0
    # no column meets the mesh: a single class -> 1, then latched 2
    xy_far = np.ascontiguousarray(xy + 100.0)
    assert lib.fe_mesh_create(d(corners), len(corners), d(anchor), d(xy_far), *shape, 1.0, d(zs), 64,
                              ct.byref(h)) == 0
    assert lib.fe_mesh_query(h, *map(d, out)) == 1
    assert lib.fe_mesh_query(h, *map(d, out)) == 2
    assert lib.fe_mesh_destroy(ct.byref(h)) == 0
    # a winding point on a mesh corner -> 5
    v, f = arrays["capacity_vertices"], arrays["capacity_faces"]
    pts = np.ascontiguousarray(np.concatenate([v[:1], np.zeros((7, 3))]))
    assert lib.fe_gwn_create(d(np.ascontiguousarray(v[f])), len(f), d(pts), 2, 2, 2, 1.0, ct.byref(h)) == 0
    assert lib.fe_mesh_query(h, *[d(np.empty((2, 2, 2), t)) for t in (np.uint8, np.uint8, np.float32)]) == 5
    assert lib.fe_mesh_destroy(ct.byref(h)) == 0


def test_device_replays(fe, gpu):
    """Replays on torch's stream, with eager work between them, leave the
    same results as a query."""
    import torch
    name, v, f, pitch, lo = generated_cases()[0]
    with fe.PreparedMeshField(v, f, pitch, lo) as field:
        gmin, shape, surface, solid, distance = field.evaluate()
        x = torch.randn(1 << 20, device=gpu)
        for _ in range(3):
            field.run_device(torch.cuda.current_stream())
            x = x * 1.0001 + 1.0
            torch.cuda.synchronize()
            assert field.device_status() == 0
        again = field.evaluate()
        assert np.array_equal(again[3], solid) and again[4].tobytes() == distance.tobytes()
        ptrs = field.device_buffers()
        assert all(p for p in ptrs[:7]) and ptrs[7] == 64


def test_result_region_reference_case(fe, gpu, fx):
    """A prepared stage's query written into a region of the service layout."""
    meta, arrays = fx
    v, f, pitch, lo, _ = reference_case(meta, arrays, "rotated_0.0")
    want = meta["cases"]["rotated_0.0"]["digests"]
    field = fe.PreparedMeshField(v, f, pitch, lo)
    try:
        region = bytearray(fe.mesh_service.layout(field.shape)[2])
        for leg in range(2):
            solid, surface, distance = fe.mesh_service.fill_region(field, region)
            got = [digest(x) for x in (surface, solid, distance)]
            assert got == [want["surface"], want["solid"], want["distance"]], leg
    finally:
        field.close()
