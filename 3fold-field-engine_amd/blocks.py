"""Leaf-box packed fields (lossless adaptive sample blocks) on the gfx950 library.

Drop-in for FE/adaptive_sample_blocks_v1.py:61-151 (AdaptiveSampleBlocks) and
FE/adaptive_pack_native_v1.py:11-61 (NativeAdaptiveBlocks): same constructor
validation and errors, the same `nodes` table (52-byte NODE records,
pre-order) and `payload`, and the same `samples_at` / `query` results.  The
packing, lookups and trilinear queries run on the device; a field also keeps
device copies of its table and payload for the CSG executor.
"""
import ctypes as ct
import operator
from numbers import Real

import numpy as np

from . import _lib

NODE = np.dtype([("lo", "<i4", (3,)), ("shape", "<i4", (3,)), ("stored_shape", "<i4", (3,)),
                 ("children", "<i4", (2,)), ("offset", "<i8")])
assert NODE.itemsize == 52


def _torch():
    import torch
    return torch


def _validate(samples, origin, pitch, leaf_samples):
    """Constructor checks of FE/adaptive_sample_blocks_v1.py:62-86."""
    if isinstance(samples, np.ndarray) or not hasattr(samples, "is_cuda"):
        samples = np.asarray(samples)
        if samples.dtype != np.float32 or samples.ndim != 3 or not samples.size:
            raise ValueError("Nonempty three-dimensional float32 samples required")
        if max(samples.shape) > 4096 or samples.size > 16777216 or not np.isfinite(samples).all():
            raise ValueError("Finite bounded field required")
    origin = np.asarray(origin, dtype=np.float64)
    if origin.shape != (3,) or not np.isfinite(origin).all():
        raise ValueError("Finite three-coordinate origin required")
    if isinstance(pitch, (bool, np.bool_)) or not isinstance(pitch, Real):
        raise ValueError("Positive finite scalar pitch required")
    try:
        pitch = float(pitch)
    except (ValueError, TypeError):
        raise ValueError("Positive finite scalar pitch required") from None
    if not np.isfinite(pitch) or pitch <= 0:
        raise ValueError("Positive finite scalar pitch required")
    if isinstance(leaf_samples, (bool, np.bool_)):
        raise ValueError("Integer leaf size required")
    try:
        leaf_samples = operator.index(leaf_samples)
    except TypeError:
        raise ValueError("Integer leaf size required") from None
    if not 8 <= leaf_samples <= 4096:
        raise ValueError("Leaf size outside bounded contract")
    return samples, origin, pitch, leaf_samples


def pack_device(dense, leaf_samples=512, stream=None):
    """Packs a device float32 (nx,ny,nz) tensor; returns device (nodes uint8 (n,52), payload float32)."""
    torch = _torch()
    L = _lib.lib()
    status = ct.c_int(0)
    nx, ny, nz = dense.shape
    h = L.fe_pack_device(_lib.ptr(dense), nx, ny, nz, int(leaf_samples), ct.byref(status), _lib.stream_handle(stream))
    if not h:
        if status.value == 1:
            raise ValueError("Finite bounded field required")
        raise RuntimeError("Device packing failed: " + L.fe_last_error().decode())
    try:
        n, m = L.fe_pack_nodes(h), L.fe_pack_samples(h)
        nodes = torch.empty((n, 52), dtype=torch.uint8, device=dense.device)
        payload = torch.empty(max(m, 1), dtype=torch.float32, device=dense.device)
        if L.fe_pack_copy(h, _lib.ptr(nodes), n, _lib.ptr(payload), m, _lib.stream_handle(stream)):
            raise RuntimeError("Device packing copy failed")
    finally:
        L.fe_pack_free(h)
    return nodes, payload[:m]


class AdaptiveSampleBlocks:
    """FE/adaptive_sample_blocks_v1.py:61-151 on the device.

    `samples` may be a host float32 array or a device float32 tensor (the CSG
    executor passes its output without a host round trip)."""

    def __init__(self, samples, origin, pitch, *, leaf_samples=512):
        torch = _torch()
        samples, origin, pitch, leaf_samples = _validate(samples, origin, pitch, leaf_samples)
        if isinstance(samples, np.ndarray):
            dense = torch.from_numpy(np.ascontiguousarray(samples)).to("cuda:0")
        else:
            if samples.dtype != torch.float32 or samples.dim() != 3 or samples.numel() == 0:
                raise ValueError("Nonempty three-dimensional float32 samples required")
            # a device field keeps the format's 4096-per-axis bound only; the
            # reference's 2^24-sample cap applies to host arrays (drop-in path)
            if max(samples.shape) > 4096:
                raise ValueError("Finite bounded field required")
            dense = samples.contiguous()
        self.d_nodes, self.d_payload = pack_device(dense, leaf_samples)
        self.nodes = self.d_nodes.cpu().numpy().reshape(-1).view(NODE)
        self.payload = self.d_payload.cpu().numpy()
        self.origin = origin.copy()
        self.pitch = pitch
        self.shape = tuple(int(x) for x in dense.shape)
        for array in (self.nodes, self.payload, self.origin):
            array.flags.writeable = False
        self._dev_key = (id(self.nodes), id(self.payload))

    @classmethod
    def from_arrays(cls, nodes, payload, origin, pitch, shape):
        """A field from an existing node table and payload (uploaded once)."""
        torch = _torch()
        self = cls.__new__(cls)
        self.nodes = np.ascontiguousarray(nodes, NODE)
        self.payload = np.ascontiguousarray(payload, np.float32)
        self.origin = np.asarray(origin, np.float64).copy()
        self.pitch = float(pitch)
        self.shape = tuple(int(x) for x in shape)
        self._device()
        return self

    def _device(self):
        """Device copies of the table and payload (uploaded again when the host
        arrays were replaced)."""
        torch = _torch()
        key = (id(self.nodes), id(self.payload))
        if getattr(self, "_dev_key", None) != key:
            self.d_nodes = torch.from_numpy(np.ascontiguousarray(self.nodes, NODE).view(np.uint8)
                                            .reshape(-1, 52).copy()).to("cuda:0")
            self.d_payload = torch.from_numpy(np.ascontiguousarray(self.payload, np.float32).copy()).to("cuda:0")
            self._dev_key = key
        return self.d_nodes, self.d_payload

    @property
    def storage_bytes(self):
        return self.nodes.nbytes + self.payload.nbytes + self.origin.nbytes + 8 + 3 * 8

    def samples_at(self, indices):
        torch = _torch()
        indices = np.asarray(indices)
        if indices.ndim != 2 or indices.shape[1] != 3 or not np.issubdtype(indices.dtype, np.integer):
            raise ValueError("Integer index triples required")
        if np.any(indices < 0) or np.any(indices >= np.array(self.shape)):
            raise ValueError("Sample index outside field")
        nodes, payload = self._device()
        idx = torch.from_numpy(np.ascontiguousarray(indices, np.int64)).to(nodes.device)
        out = torch.empty(len(indices), dtype=torch.int32, device=nodes.device)
        if len(indices):
            status = _lib.lib().fe_samples_at(_lib.ptr(nodes), _lib.ptr(payload), _lib.ptr(idx), len(indices),
                                              _lib.ptr(out), _lib.stream_handle(None))
            if status:
                raise RuntimeError(f"samples_at status {status}")
        return out.cpu().numpy().view(np.float32)

    def query(self, points):
        torch = _torch()
        points = np.asarray(points, dtype=np.float64)
        if points.ndim != 2 or points.shape[1] != 3 or not np.isfinite(points).all():
            raise ValueError("Finite point triples required")
        with np.errstate(over="ignore", invalid="ignore"):
            coordinates = (points - self.origin) / self.pitch
        if not np.isfinite(coordinates).all():
            raise ValueError("Query exceeds finite grid-coordinate range")
        nodes, payload = self._device()
        pts = torch.from_numpy(np.ascontiguousarray(points)).to(nodes.device)
        org = torch.from_numpy(self.origin.copy()).to(nodes.device)
        out = torch.empty(len(points), dtype=torch.float32, device=nodes.device)
        if len(points):
            status = _lib.lib().fe_query_packed(_lib.ptr(nodes), _lib.ptr(payload), *self.shape,
                                                self.origin.ctypes.data_as(ct.POINTER(ct.c_double)), self.pitch,
                                                _lib.ptr(pts), len(points), _lib.ptr(out), _lib.stream_handle(None))
            if status:
                raise RuntimeError(f"query status {status}")
        del org
        return out.cpu().numpy()

    def dense_device(self):
        """The field as a dense device float32 tensor."""
        torch = _torch()
        nodes, payload = self._device()
        out = torch.empty(self.shape, dtype=torch.float32, device=nodes.device)
        status = _lib.lib().fe_unpack(_lib.ptr(nodes), len(self.nodes), _lib.ptr(payload), *self.shape, _lib.ptr(out),
                                      _lib.stream_handle(None))
        if status:
            raise RuntimeError(f"unpack status {status}")
        return out


class NativeAdaptiveBlocks(AdaptiveSampleBlocks):
    """FE/adaptive_pack_native_v1.py:11-61: `library` names a library exporting
    adaptive_pack/adaptive_nodes/adaptive_samples/adaptive_copy/adaptive_free
    (this package's libfe_hip.so by default); the packing runs there."""

    def __init__(self, samples, origin, pitch, *, library=None, leaf_samples=512):
        samples, origin, pitch, leaf_samples = _validate(samples, origin, pitch, leaf_samples)
        if not isinstance(samples, np.ndarray):
            samples = samples.cpu().numpy()
        L = _lib.lib() if library is None else ct.CDLL(str(library))
        L.adaptive_pack.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int]
        L.adaptive_pack.restype = ct.c_void_p
        for name in ("adaptive_nodes", "adaptive_samples"):
            fn = getattr(L, name)
            fn.argtypes = [ct.c_void_p]
            fn.restype = ct.c_int64
        L.adaptive_copy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64]
        L.adaptive_copy.restype = ct.c_int
        L.adaptive_free.argtypes = [ct.c_void_p]
        L.adaptive_free.restype = None
        array = np.ascontiguousarray(samples)
        shape = np.array(array.shape, dtype=np.int32)
        pointer = L.adaptive_pack(array.ctypes.data, shape.ctypes.data, leaf_samples)
        if not pointer:
            raise ValueError("Native packing failed")
        try:
            nodes = np.empty(L.adaptive_nodes(pointer), dtype=NODE)
            payload = np.empty(L.adaptive_samples(pointer), dtype=np.float32)
            if L.adaptive_copy(pointer, nodes.ctypes.data, len(nodes), payload.ctypes.data, len(payload)):
                raise RuntimeError("Native packing copy failed")
        finally:
            L.adaptive_free(pointer)
        self.nodes, self.payload, self.origin, self.pitch, self.shape = nodes, payload, origin.copy(), pitch, array.shape
        for value in (self.nodes, self.payload, self.origin):
            value.flags.writeable = False
