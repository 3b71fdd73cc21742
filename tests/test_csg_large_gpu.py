"""The CSG executor beyond its LDS slots and beyond the reference's sample
cap, the check-once operand path, and packing past 2^24 samples.

* A 256-field, 1024-instruction program that keeps about 100 values live
  (more than the 16 LDS slots, so values spill to the device arena) equals
  the oracle's csg_execute (oracle/fe_oracle.c, the restatement of
  FE/csg_batch_native_v1/execute.cpp:10-54) and the reference's own
  csg_execute compiled from source (oracle/_ref), through every entry point;
  status codes of spilling programs (non-finite results, malformed and
  coverage faults at chosen tiles) equal theirs.
* fe_csg_check's first failing tile equals the first reference tile whose
  gather (oracle or_gather, FE/csg_tile_native_v1/gather.cpp:11-50) fails.
* Past 2^24 samples (where the reference refuses) the device executor equals
  the dense restatement of FE/field_csg_tiled_v1.py:84-89 below, itself
  pinned to the oracle's csg_execute on small grids.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def encode(prog):
    P = np.zeros(len(prog), O.INSTR)
    for i, (op, l, r) in enumerate(prog):
        P[i]["op"] = {"union": 0, "difference": 1, "offset": 2}[op]
        P[i]["left"] = l
        if op == "offset":
            P[i]["offset"] = r
        else:
            P[i]["right"] = r
    return P


def wide_program(nf=256, n=1024, live=100, seed=0):
    """`live` values v_k = union(f_k, f_k+live) first, then a chain of
    differences, unions and offsets that draws on every v_k again and again
    (so all of them stay live to the end)."""
    rng = np.random.default_rng(seed)
    prog = [("union", k, k + live) for k in range(live)]
    acc = nf
    while len(prog) < n:
        i = nf + len(prog) - 1
        v = nf + int(rng.integers(0, live))
        r = rng.random()
        if r < 0.45:
            prog.append(("difference", i, v))
        elif r < 0.9:
            prog.append(("union", i, v))
        else:
            prog.append(("offset", i, float(rng.uniform(-0.5, 0.5))))
        acc = i
    return prog


def fields_of(rng, nf, shape):
    out = []
    for k in range(nf):
        a = rng.normal(size=shape).astype(np.float32)
        if k % 3 == 0:
            a = np.repeat(a[:, :, :1], shape[2], axis=2)  # z-invariant leaves
        out.append(a)
    return out


def run_host(lib, packed, prog, shape):
    nodes = (ct.c_void_p * len(packed))(*[p[0].ctypes.data for p in packed])
    pays = (ct.c_void_p * len(packed))(*[p[1].ctypes.data for p in packed])
    counts = np.array([len(p[0]) for p in packed], np.int64)
    lens = np.array([len(p[1]) for p in packed], np.int64)
    sh = np.array(shape, np.int32)
    out = np.empty(shape, np.float32)
    code = lib.csg_execute(nodes, counts.ctypes.data, pays, lens.ctypes.data, len(packed), sh.ctypes.data,
                           prog.ctypes.data, len(prog), out.ctypes.data)
    return code, out


def ref_csg():
    ref = O.ref_lib("csg_batch")
    if ref is not None:
        ref.csg_execute.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                    ct.c_void_p, ct.c_int32, ct.c_void_p]
        ref.csg_execute.restype = ct.c_int
    return ref


def test_wide_program_spills_and_matches(fe, gpu):
    import torch
    rng = np.random.default_rng(21)
    shape = (30, 21, 40)
    arrays = fields_of(rng, 256, shape)
    prog = wide_program()
    enc = encode(prog)
    packed = [O.pack(a) for a in arrays]
    code, want = O.csg_execute(packed, enc, shape)
    assert code == 0
    ref = ref_csg()
    if ref is not None:
        rcode, rout = run_host(ref, packed, enc, shape)
        assert rcode == 0 and rout.tobytes() == want.tobytes()
    lib = fe._lib.lib()
    sh = np.array(shape, np.int32)
    ws = lib.fe_csg_workspace_bytes(256, sh.ctypes.data, enc.ctypes.data, len(enc))
    print("workspace bytes", ws, flush=True)
    assert ws > 4 << 20  # the slots beyond LDS live in the arena
    fields = [fe.AdaptiveSampleBlocks(a, [0, 0, 0], 1) for a in arrays]
    _, got = fe.csg.evaluate_dense(fields, prog)
    assert got.cpu().numpy().tobytes() == want.tobytes()
    _, again = fe.csg.evaluate_dense(fields, prog)  # cached checks and workspace
    assert torch.equal(again, got)
    hcode, hout = run_host(lib, packed, enc, shape)
    assert hcode == 0 and hout.tobytes() == want.tobytes()
    # fe_csg_device on the device tables
    dev = [f._device() for f in fields]
    nodes = (ct.c_void_p * 256)(*[d[0].data_ptr() for d in dev])
    pays = (ct.c_void_p * 256)(*[d[1].data_ptr() for d in dev])
    counts = np.array([d[0].shape[0] for d in dev], np.int64)
    lens = np.array([d[1].numel() for d in dev], np.int64)
    out = torch.empty(shape, dtype=torch.float32, device=gpu)
    assert lib.fe_csg_device(nodes, counts.ctypes.data, pays, lens.ctypes.data, 256, sh.ctypes.data, enc.ctypes.data,
                             len(enc), fe._lib.ptr(out), fe._lib.stream_handle(None)) == 0
    assert out.cpu().numpy().tobytes() == want.tobytes()


def test_spilling_program_status_codes(fe, gpu):
    rng = np.random.default_rng(22)
    shape = (40, 50, 30)
    arrays = fields_of(rng, 256, shape)
    arrays[7] = (rng.normal(size=shape) * 1e37).astype(np.float32)
    base = wide_program(seed=3)
    packed = [O.pack(a) for a in arrays]
    lib = fe._lib.lib()
    ref = ref_csg()
    variants = {"ok": base,
                # values of field 7 (around 1e37) pushed past the float range
                # mid-program (where they reach the chain) and at the end
                "overflow_mid": base[:600] + [("offset", 256 + 599, -3e38)] + base[601:],
                "overflow_end": base[:-1] + [("offset", 7, -3.5e38)]}
    n0, p0 = packed[5]
    leaf = int(np.argmax(n0["children"][:, 0] < 0))
    last = len(n0) - 1 - int(np.argmax(n0["children"][::-1, 0] < 0))
    faults = {"none": n0}
    gap = n0.copy()
    gap["shape"][last] = [1, 1, 1]
    gap["stored_shape"][last] = [1, 1, 1]
    faults["gap_late"] = gap
    bad = n0.copy()
    bad["offset"][leaf] = 10 ** 9
    faults["malformed"] = bad
    faults["overlap"] = np.concatenate((n0, n0[leaf:leaf + 1]))
    for fname, nodes in faults.items():
        ops = [p for p in packed]
        ops[5] = (nodes, p0)
        for vname, prog in variants.items():
            enc = encode(prog)
            want, _ = O.csg_execute(ops, enc, shape)
            got, _ = run_host(lib, ops, enc, shape)
            print("status", fname, vname, want, got, flush=True)
            assert got == want, (fname, vname)
            if ref is not None:
                assert run_host(ref, ops, enc, shape)[0] == want, (fname, vname)


def first_failing_tile(nodes, payload, shape):
    k = 0
    for x in range(0, shape[0], 16):
        for y in range(0, shape[1], 32):
            for z in range(0, shape[2], 16):
                lo = np.array([x, y, z], np.int32)
                hi = np.minimum(lo + [16, 32, 16], shape).astype(np.int32)
                if O.gather(nodes, payload, lo, hi)[0]:
                    return k
                k += 1
    return 2 ** 31 - 1


def test_operand_check_first_tile(fe, gpu):
    import torch
    rng = np.random.default_rng(23)
    shape = (37, 70, 33)
    a = rng.normal(size=shape).astype(np.float32)
    a[:, :40] = 1.5  # large constant leaves next to small ones
    n0, p0 = O.pack(a, leaf_samples=64)
    leaves = np.flatnonzero(n0["children"][:, 0] < 0)
    cases = {"valid": n0}
    for name, pick in (("gap_first", leaves[0]), ("gap_mid", leaves[len(leaves) // 2]), ("gap_last", leaves[-1])):
        t = n0.copy()
        t["shape"][pick] = np.maximum(t["shape"][pick] // 2, 1)
        t["stored_shape"][pick] = np.where(t["stored_shape"][pick] == 1, 1, t["shape"][pick])
        cases[name] = t
    t = n0.copy()
    t["lo"][leaves[len(leaves) // 3]][2] += 1  # shifted: one overlap row and one gap row
    cases["shifted"] = t
    t = n0.copy()
    t["stored_shape"][leaves[3]] = t["shape"][leaves[3]] + 1
    cases["malformed"] = t
    L = fe._lib.lib()
    sh = np.array(shape, np.int32)
    pd = torch.from_numpy(p0.view(np.uint32).view(np.float32).copy()).to(gpu)
    for name, nodes in cases.items():
        want = 0 if name == "malformed" else first_failing_tile(nodes, p0, shape)
        nd = torch.from_numpy(nodes.view(np.uint8).copy()).to(gpu)
        got = ct.c_int32(-5)
        assert L.fe_csg_check(fe._lib.ptr(nd), len(nodes), fe._lib.ptr(pd), len(p0), sh.ctypes.data, ct.byref(got),
                              fe._lib.stream_handle(None)) == 0
        print("check", name, want, got.value, flush=True)
        assert got.value == want, name


def dense_program(arrays, prog):
    """Dense restatement of FE/field_csg_tiled_v1.py:84-89: union a<b?a:b,
    difference a>(-b)?a:(-b), offset f32(f64(a) - d), in program order."""
    vals = list(arrays)
    for op, l, r in prog:
        a = vals[l]
        if op == "offset":
            vals.append((a.astype(np.float64) - r).astype(np.float32))
        else:
            b = vals[r] if op == "union" else -vals[r]
            vals.append(np.where(a < b, a, b) if op == "union" else np.where(a > b, a, b))
    return vals[-1]


def test_dense_restatement_pinned_small():
    rng = np.random.default_rng(24)
    shape = (19, 35, 21)
    arrays = [rng.normal(size=shape).astype(np.float32) for _ in range(4)]
    arrays[1].ravel()[::3] = np.float32(-0.0)
    arrays[2].ravel()[::5] = np.float32(0.0)
    prog = [("union", 0, 1), ("difference", 4, 2), ("offset", 5, 0.25), ("union", 3, 6), ("difference", 1, 2)]
    code, want = O.csg_execute([O.pack(a) for a in arrays], encode(prog), shape)
    assert code == 0
    assert dense_program(arrays, prog).tobytes() == want.tobytes()


def test_beyond_sample_cap(fe, gpu):
    """272 x 256 x 256 = 17.8 M samples: spheres and boxes packed from device
    tensors, one program, against the dense restatement."""
    import torch
    shape = (272, 256, 256)
    g = [torch.arange(n, dtype=torch.float32, device=gpu) for n in shape]
    X, Y, Z = torch.meshgrid(*g, indexing="ij")
    s1 = torch.sqrt((X - 100) ** 2 + (Y - 120) ** 2 + (Z - 128) ** 2) - 70
    s2 = torch.sqrt((X - 170) ** 2 + (Y - 130) ** 2 + (Z - 120) ** 2) - 60
    qx, qy, qz = (X - 136).abs() - 90, (Y - 128).abs() - 20, (Z - 128).abs() - 100
    box = torch.sqrt(qx.clamp(min=0) ** 2 + qy.clamp(min=0) ** 2 + qz.clamp(min=0) ** 2) + \
        torch.clamp(torch.maximum(qx, torch.maximum(qy, qz)), max=0)
    dev = [s1.contiguous(), s2.contiguous(), box.contiguous()]
    fields = [fe.AdaptiveSampleBlocks(t, [0, 0, 0], 1) for t in dev]
    prog = [("union", 0, 1), ("difference", 3, 2), ("offset", 4, 1.5)]
    _, got = fe.csg.evaluate_dense(fields, prog)
    want = dense_program([t.cpu().numpy() for t in dev], prog)
    assert got.cpu().numpy().tobytes() == want.tobytes()
    # the packs are lossless
    for f, t in zip(fields, dev):
        assert torch.equal(f.dense_device(), t)
