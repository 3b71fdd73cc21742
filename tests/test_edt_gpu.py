"""GPU parity of the exact integer EDT against the reference's golden data and the oracle."""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unpack(bits, shape):
    return np.unpackbits(bits)[:int(np.prod(shape))].reshape(shape).astype(bool)


def test_integer_families_through_abi(fe, golden, gpu):
    from conftest import ROOT
    z = np.load(ROOT / "tests/golden/edt_integer.npz")
    edt = fe.NativeEDT()
    for key, meta in golden["edt_integer"].items():
        shape = tuple(meta["shape"])
        mask = unpack(z[f"{key}_mask"], shape)
        print("case", meta["name"], shape, flush=True)
        s0 = edt.squared_distance(~mask)
        s1 = edt.squared_distance(mask)
        assert np.array_equal(s0, z[f"{key}_sq0"]) and np.array_equal(s1, z[f"{key}_sq1"]), meta["name"]
        assert np.array_equal(fe.squared_distance(mask), z[f"{key}_sq1"].astype(np.int64))
        for pitch, digest in meta["signed"].items():
            assert sha(edt.signed_distance(mask, float(pitch))) == digest, (meta["name"], pitch)
            assert sha(fe.signed_distance(mask, float(pitch))) == digest


def test_pair_cases_persistent(fe, golden, gpu):
    from conftest import ROOT
    z = np.load(ROOT / "tests/golden/edt_pairs.npz")
    for key, meta in golden["edt_pairs"].items():
        shape = tuple(meta["shape"])
        mask = unpack(z[f"{key}_mask"], shape)
        print("case", meta["source"], meta["name"], shape, flush=True)
        with fe.PersistentPairEDT(shape) as h:
            for pitch, digest in meta["signed"].items():
                out, sq = h.query(mask, float(pitch), with_squared=True)
                assert sha(sq[0]) == meta["squared0"] and sha(sq[1]) == meta["squared1"], meta["name"]
                assert sha(out) == digest, (meta["name"], pitch)
        large = fe.NativeEDTLarge()
        assert sha(large.squared_distance(~mask)) == meta["squared0"]


def random_masks(rng):
    for shape in ((1, 1, 2), (3, 70, 5), (130, 9, 11), (17, 300, 3), (64, 64, 64), (5, 6, 1100), (1100, 3, 4),
                  (2, 1030, 7), (4, 5, 16), (20, 30, 17), (9, 8, 31), (3, 2, 4096),
                  # every axis <= 1024 with lines of 513..1024 positions (48 and 64 chunks per line)
                  (3, 900, 5), (1000, 4, 6), (600, 7, 9), (5, 1000, 17), (8, 700, 40)):
        yield shape, rng.random(shape) < rng.uniform(0.01, 0.9)
    m = np.zeros((90, 80, 70), bool)
    m[45, 40, 35] = True
    yield m.shape, m
    yield m.shape, ~m


def test_random_shapes_vs_oracle(fe, gpu):
    import torch
    rng = np.random.default_rng(21)
    for shape, mask in random_masks(rng):
        if mask.all() or not mask.any():
            continue
        print("shape", shape, flush=True)
        d = torch.from_numpy(mask.astype(np.uint8)).to(gpu)
        pair = fe.edt_device(d, planes=2).cpu().numpy()
        _, sq = O.edt_pair(mask, 1.0)
        assert np.array_equal(pair, sq), shape
        one = fe.edt_device(d, planes=1).cpu().numpy()
        assert np.array_equal(one, O.edt_envelope(mask)), shape
        signed = fe.signed_from_pair_device(torch.from_numpy(sq).to(gpu), 0.3).cpu().numpy()
        assert signed.tobytes() == O.edt_signed_from_squares(sq[0], sq[1], 0.3).tobytes()


def test_errors_and_latch(fe, gpu):
    edt = fe.NativeEDTLarge()
    with pytest.raises(ValueError):
        edt.squared_distance(np.zeros((4, 4, 4), bool))
    with pytest.raises(ValueError):
        edt.squared_distance(np.zeros((513, 2, 2), bool))
    h = fe.PersistentPairEDT((4, 5, 6))
    m = np.zeros((4, 5, 6), bool)
    m[1, 2, 3] = True
    with pytest.raises(ValueError):
        h.query(m, float("nan"))
    with pytest.raises(ValueError):
        h.query(np.zeros((4, 5, 7), bool) | m[:, :, :1].repeat(7, 2), 1.0)
    h.close()
    with pytest.raises(RuntimeError):
        h.query(m, 1.0)
    import threading
    h2 = fe.PersistentPairEDT((4, 5, 6))
    err = []
    t = threading.Thread(target=lambda: err.append(pytest.raises(RuntimeError, h2.query, m, 1.0)))
    t.start()
    t.join()
    assert err
    h2.close()
