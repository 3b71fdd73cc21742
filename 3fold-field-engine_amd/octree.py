"""2-bit adaptive block tree on the gfx950 library.

Drop-in for FE/field_adaptive_blocks_v1.py:20-138: `AdaptiveBlocks` keeps the
reference's fields (shape, pitch, root_shape, root_count, node_count, packed,
branch_prefix, value_prefix, values) and methods (kinds, rank_before, sample,
stored_bytes, fingerprint); `build(wp, dense, pitch)` returns the same
(tree, histogram) pair (`wp` is accepted and ignored).  The dense-source
classification at block sides 8, 4, 2, 1, the breadth-first construction,
the packing and the rank checkpoints run on the device, and so do the
lookups of `sample`.
"""
import ctypes as ct
import hashlib

import numpy as np

from . import _lib

LOOKUP = np.array([[sum(((byte >> (2 * i)) & 3) == code for i in range(4)) for byte in range(256)]
                   for code in range(4)], dtype=np.int32)
SIDES = (8, 4, 2, 1)


def _torch():
    import torch
    return torch


class AdaptiveBlocks:
    """FE/field_adaptive_blocks_v1.py:20-94."""

    def __init__(self, shape, pitch, root_shape, root_count, node_count, packed, branch_prefix, value_prefix, values):
        self.shape, self.pitch = tuple(int(x) for x in shape), float(pitch)
        self.root_shape, self.root_count, self.node_count = tuple(int(x) for x in root_shape), int(root_count), \
            int(node_count)
        self.packed = np.ascontiguousarray(packed, np.uint8)
        self.branch_prefix = np.ascontiguousarray(branch_prefix, np.int32)
        self.value_prefix = np.ascontiguousarray(value_prefix, np.int32)
        self.values = np.ascontiguousarray(values, np.float32)
        self._dev = None

    def kinds(self, nodes):
        nodes = np.asarray(nodes)
        return (self.packed[nodes // 4] >> (2 * (nodes % 4))) & 3

    def rank_before(self, nodes, code):
        """Nodes before `nodes` with `code` (2 or 3): checkpoint of the 64-byte
        group plus the codes between it and the node (FE/field_adaptive_blocks_v1.py:35-50)."""
        nodes = np.asarray(nodes, np.int64)
        byte = nodes // 4
        group = byte // 64
        prefix = self.branch_prefix if code == 2 else self.value_prefix
        rank = prefix[group].astype(np.int64)
        for offset in range(63):
            b = group * 64 + offset
            take = b < byte
            if not take.any():
                break
            rank[take] += LOOKUP[code, self.packed[b[take]]]
        partial = self.packed[byte]
        for slot in range(3):
            rank += (nodes % 4 > slot) & (((partial >> (2 * slot)) & 3) == code)
        return rank

    def _device(self):
        torch = _torch()
        if self._dev is None:
            dev = torch.device("cuda:0")
            padded = np.zeros(len(self.branch_prefix) * 64, np.uint8)
            padded[:len(self.packed)] = self.packed
            vals = self.values if len(self.values) else np.zeros(1, np.float32)
            self._dev = (torch.from_numpy(padded).to(dev), torch.from_numpy(self.branch_prefix.copy()).to(dev),
                         torch.from_numpy(self.value_prefix.copy()).to(dev), torch.from_numpy(vals.copy()).to(dev))
        return self._dev

    def sample(self, indices):
        """Values at integer grid coordinates [N,3] (FE/field_adaptive_blocks_v1.py:52-82)."""
        torch = _torch()
        ijk = np.asarray(indices, dtype=np.int64)
        if ijk.ndim != 2 or ijk.shape[1] != 3 or np.any(ijk < 0) or np.any(ijk >= self.shape):
            raise ValueError("Expected in-domain integer grid coordinates[N,3]")
        packed, bpre, vpre, vals = self._device()
        rs = np.ascontiguousarray(self.root_shape, np.int32)
        q = torch.from_numpy(np.ascontiguousarray(ijk)).to(packed.device)
        out = torch.empty(len(ijk), dtype=torch.float32, device=packed.device)
        flag = torch.zeros(1, dtype=torch.int32, device=packed.device)
        status = _lib.lib().fe_octree_sample(_lib.ptr(packed), _lib.ptr(bpre), _lib.ptr(vpre), _lib.ptr(vals),
                                             rs.ctypes.data, self.root_count, self.pitch, _lib.ptr(q), len(ijk),
                                             _lib.ptr(out), _lib.ptr(flag), _lib.stream_handle(None))
        if status:
            raise RuntimeError(f"octree sample status {status}")
        if int(flag.item()):
            raise RuntimeError("Invalid tree depth")
        return out.cpu().numpy()

    def stored_bytes(self):
        # fixed numeric header (shapes, counts, pitch, root size) plus the arrays
        return 80 + sum(a.nbytes for a in (self.packed, self.branch_prefix, self.value_prefix, self.values))

    def fingerprint(self):
        h = hashlib.sha256()
        h.update(np.array([*self.shape, *self.root_shape, self.root_count, self.node_count], np.int64).tobytes())
        h.update(np.float64(self.pitch).tobytes())
        for a in (self.packed, self.branch_prefix, self.value_prefix, self.values):
            h.update(a.tobytes())
        return h.hexdigest()


def build(wp, dense, pitch):
    """FE/field_adaptive_blocks_v1.py:97-138 (`wp` is ignored): returns
    (AdaptiveBlocks, histogram) with histogram labels side{s}_kind{code} of the
    leaves in first-occurrence (breadth-first) order."""
    torch = _torch()
    if hasattr(dense, "is_cuda"):
        if dense.dtype != torch.float32 or dense.dim() != 3:
            raise ValueError("Three-dimensional float32 field required")
        d = dense.contiguous()
    else:
        d = torch.from_numpy(np.ascontiguousarray(dense, np.float32)).to("cuda:0")
    if d.dim() != 3 or min(d.shape) < 1 or max(d.shape) > 4096:
        raise ValueError("Three-dimensional field with axes in [1, 4096] required")
    nx, ny, nz = (int(x) for x in d.shape)
    L = _lib.lib()
    st = _lib.stream_handle(None)
    kinds = []
    for side in SIDES:
        nb = (-(-nx // side)) * (-(-ny // side)) * (-(-nz // side))
        k = torch.empty(nb, dtype=torch.int32, device=d.device)
        margin = float(np.float32(np.sqrt(3.0) * side * float(pitch)))
        if L.fe_dense_classify(_lib.ptr(d), nx, ny, nz, side, margin, _lib.ptr(k), st):
            raise RuntimeError("dense classify failed")
        kinds.append(k)
    ptrs = (ct.c_void_p * 4)(*[_lib.ptr(k) for k in kinds])
    status = ct.c_int(0)
    h = L.fe_octree_build(_lib.ptr(d), nx, ny, nz, ptrs, ct.byref(status), st)
    if not h:
        raise RuntimeError(f"octree build status {status.value}: " + L.fe_last_error().decode())
    try:
        sizes = np.zeros(4, np.int64)
        counts = np.zeros(16, np.int64)
        first = np.zeros(16, np.int64)
        L.fe_octree_info(h, sizes.ctypes.data, counts.ctypes.data, first.ctypes.data)
        nodes, nvalues, nbytes, groups = (int(x) for x in sizes)
        packed = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=d.device)
        bpre = torch.empty(groups, dtype=torch.int32, device=d.device)
        vpre = torch.empty(groups, dtype=torch.int32, device=d.device)
        vals = torch.empty(max(nvalues, 1), dtype=torch.float32, device=d.device)
        if L.fe_octree_copy(h, _lib.ptr(packed), _lib.ptr(bpre), _lib.ptr(vpre), _lib.ptr(vals), st):
            raise RuntimeError("octree copy failed")
        torch.cuda.synchronize()
    finally:
        L.fe_octree_free(h)
    root_shape = tuple(-(-n // 8) for n in (nx, ny, nz))
    tree = AdaptiveBlocks((nx, ny, nz), float(pitch), root_shape, int(np.prod(root_shape)), nodes,
                          packed.cpu().numpy()[:nbytes], bpre.cpu().numpy(), vpre.cpu().numpy(),
                          vals.cpu().numpy()[:nvalues])
    counts, first = counts.reshape(4, 4), first.reshape(4, 4)
    histogram = {}
    for level, side in enumerate(SIDES):
        for code in sorted((0, 1, 3), key=lambda c: (first[level, c] < 0, first[level, c])):
            if counts[level, code]:
                histogram[f"side{side}_kind{code}"] = int(counts[level, code])
    return tree, histogram
