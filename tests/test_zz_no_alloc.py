"""fe_edt_device_static (the EDT entry point that allocates nothing) against
fe_edt_device through the C ABI: two calls in a row, the second smaller, in
both plane counts, bit for bit; sizes beyond its static scratch are refused."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _masks():
    rng = np.random.default_rng(77)
    big = (rng.random((96, 101, 103)) < 0.2).astype(np.uint8)   # 998,688 voxels: both planes fill most of it
    small = (rng.random((17, 40, 33)) < 0.35).astype(np.uint8)
    return big, small


@pytest.mark.parametrize("planes", [2, 1])
def test_static_scratch_equals_workspace_call(fe, gpu, planes):
    import torch
    L = fe._lib.lib()
    P = fe._lib.ptr
    st = fe._lib.stream_handle(None)
    outs = []
    for mask in _masks():
        print("no-alloc EDT", planes, mask.shape, flush=True)
        m = torch.from_numpy(mask).to(gpu)
        want = fe.edt_device(m, planes=planes)
        got = torch.full_like(want, -7)
        assert L.fe_edt_device_static(P(m), *mask.shape, planes, P(got), st) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), mask.shape
        outs.append(got.cpu().numpy())
    assert outs[0].shape != outs[1].shape


def test_static_scratch_refuses_oversize(fe, gpu):
    import torch
    L = fe._lib.lib()
    P = fe._lib.ptr
    m = torch.zeros(8, dtype=torch.uint8, device=gpu)
    out = torch.zeros(8, dtype=torch.int32, device=gpu)
    # 2 planes of 129*128*64 = 1,056,768 voxels exceed the 2^21-word scratch
    # (the size check returns before anything is launched)
    assert L.fe_edt_device_static(P(m), 129, 128, 64, 2, P(out), fe._lib.stream_handle(None)) == 1
    assert b"exceed" in L.fe_last_error()
    assert L.fe_edt_device_static(P(m), 0, 1, 1, 1, P(out), fe._lib.stream_handle(None)) == 1
