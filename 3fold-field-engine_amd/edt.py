"""Exact integer squared EDT and the signed field, on the gfx950 library.

Drop-in for the reference's EDT entry points, with the same names, argument
meaning, validation and error behaviour:

  squared_distance / signed_distance   FE/field_edt_integer_v1.py:14-43
  NativeEDT                            FE/native_edt_v1.py:9-46
  NativeEDTLarge                       FE/native_edt_large_v1.py:15-52
  PersistentPairEDT                    FE/edt_persistent_pair_v1.py:10-52

`library=` arguments are accepted for signature compatibility; they may name
this package's libfe_hip.so (the default) or any library exporting the same
symbols.  Device-resident entry points (`edt_pair_device`, `edt_device`)
take and return torch tensors already in HBM and are what the benchmark
times.
"""
import ctypes as ct
import operator
import threading
from numbers import Real

import numpy as np

from . import _lib


def validate(mask):
    """FE/field_edt_integer_v1.py:5-11 (<=128 per axis, <=262144 voxels)."""
    if not isinstance(mask, np.ndarray) or mask.dtype != np.bool_:
        raise ValueError('Expected a boolean array')
    if mask.ndim != 3 or min(mask.shape) < 1 or max(mask.shape) > 128 or mask.size > 262144:
        raise ValueError('Expected bounded nonempty 3D shape')
    if not mask.any() or mask.all():
        raise ValueError('Both classes must be present')


def validate_large(mask):
    """FE/native_edt_large_v1.py:6-12 (<=512 per axis, <=1048576 voxels)."""
    if not isinstance(mask, np.ndarray) or mask.dtype != np.bool_:
        raise ValueError('Expected a boolean array')
    if mask.ndim != 3 or min(mask.shape) < 1 or max(mask.shape) > 512 or mask.size > 1048576:
        raise ValueError('Expected bounded larger 3D shape')
    if not mask.any() or mask.all():
        raise ValueError('Both classes must be present')


def validate_volume(mask):
    """Uncapped contract of this library (<=4096 per axis)."""
    if not isinstance(mask, np.ndarray) or mask.dtype != np.bool_:
        raise ValueError('Expected a boolean array')
    if mask.ndim != 3 or min(mask.shape) < 1 or max(mask.shape) > 4096:
        raise ValueError('Expected nonempty 3D shape with axes <= 4096')
    if not mask.any() or mask.all():
        raise ValueError('Both classes must be present')


def _pitch(pitch):
    if isinstance(pitch, (bool, np.bool_)) or not np.isscalar(pitch):
        raise ValueError('Expected scalar isotropic pitch')
    try:
        pitch = float(pitch)
    except (ValueError, TypeError, OverflowError) as error:
        raise ValueError('Expected real pitch') from error
    if not np.isfinite(pitch) or not 1e-12 <= pitch <= 1e12:
        raise ValueError('Pitch outside bounded contract')
    return pitch


def _signed_from_pair(sq_out, sq_in, pitch):
    # Same arithmetic as FE/field_edt_integer_v1.py:40-43 (host form).
    outside = (np.sqrt(sq_out.astype(np.float64)) * pitch).astype(np.float32)
    inside = (np.sqrt(sq_in.astype(np.float64)) * pitch).astype(np.float32)
    signed = (outside - inside).astype(np.float32)
    return (signed - np.sign(signed) * np.float32(0.5 * pitch)).astype(np.float32)


def _detail(status):
    if status != 2:
        return ''
    try:
        return ': ' + _lib.lib().fe_last_error().decode()
    except Exception:
        return ''


def squared_distance(binary):
    """Squared distance from every True voxel to the nearest False voxel (int64).

    FE/field_edt_integer_v1.py:14-26 contract, computed on the GPU."""
    validate(binary)
    source = np.ascontiguousarray(binary, dtype=np.uint8)
    out = np.empty(source.shape, np.int32)
    status = _lib.lib().edt_squared(source.ctypes.data, *source.shape, out.ctypes.data)
    if status:
        raise RuntimeError(f'Native EDT status {status}' + _detail(status))
    return out.astype(np.int64)


def signed_distance(mask, pitch):
    """FE/field_edt_integer_v1.py:29-43: f32(sqrt(d_out)*h) - f32(sqrt(d_in)*h), half-pitch corrected."""
    validate(mask)
    pitch = _pitch(pitch)
    return _signed_from_pair(squared_distance(~mask), squared_distance(mask), pitch)


class _ThreadOwned:
    _label = 'EDT'

    def _check(self):
        if threading.get_ident() != self._thread:
            raise RuntimeError(f'{self._label} belongs to another thread')
        if self._failed:
            raise RuntimeError(f'{self._label} runtime previously failed; no retry')


class NativeEDT(_ThreadOwned):
    """FE/native_edt_v1.py:9-46: one host thread, bounded mixed masks (<=128 per axis)."""
    _label = 'NativeEDT'
    _symbol = 'edt_squared'
    _validate = staticmethod(validate)

    def __init__(self, library=None):
        self._thread = threading.get_ident()
        self._failed = False
        if library is None:
            self._fn = getattr(_lib.lib(), self._symbol)
        else:
            L = ct.CDLL(str(library))
            self._fn = getattr(L, self._symbol)
            self._fn.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
            self._fn.restype = ct.c_int

    def squared_distance(self, binary):
        self._check()
        self._validate(binary)
        source = np.ascontiguousarray(binary, dtype=np.uint8)
        output = np.empty(source.shape, dtype=np.int32)
        status = self._fn(source.ctypes.data, *source.shape, output.ctypes.data)
        if status:
            self._failed = status != 1
            raise RuntimeError(f'Native EDT status {status}' + _detail(status))
        return output

    def signed_distance(self, mask, pitch):
        self._check()
        self._validate(mask)
        if isinstance(pitch, (bool, np.bool_)) or not isinstance(pitch, Real):
            raise ValueError('Expected a real scalar pitch')
        pitch = float(pitch)
        if not np.isfinite(pitch) or not 1e-12 <= pitch <= 1e12:
            raise ValueError('Pitch outside bounded contract')
        return _signed_from_pair(self.squared_distance(~mask), self.squared_distance(mask), pitch)


class NativeEDTLarge(NativeEDT):
    """FE/native_edt_large_v1.py:15-52: <=512 per axis, <=1048576 voxels."""
    _label = 'NativeEDTLarge'
    _symbol = 'edt_squared_large'
    _validate = staticmethod(validate_large)


class NativeEDTVolume(NativeEDT):
    """Same contract without the reference's size caps (axes <= 4096)."""
    _label = 'NativeEDTVolume'
    _symbol = 'fe_edt_squared_host'
    _validate = staticmethod(validate_volume)


class PersistentPairEDT:
    """FE/edt_persistent_pair_v1.py:10-52: thread-owned persistent paired EDT with explicit close.

    capped=False lifts the reference's 512/2^20 size caps (axes <= 4096)."""

    def __init__(self, shape, *, library=None, capped=True):
        raw = tuple(shape)
        if len(raw) != 3 or any(isinstance(x, (bool, np.bool_)) for x in raw):
            raise ValueError('Three integer dimensions required')
        try:
            self.shape = tuple(operator.index(x) for x in raw)
        except TypeError:
            raise ValueError('Three integer dimensions required') from None
        cap = 512 if capped else 4096
        if min(self.shape) < 1 or max(self.shape) > cap or (capped and np.prod(self.shape) > 1048576):
            raise ValueError('Shape outside bounded EDT capacity')
        self._capped = capped
        self._owner = threading.get_ident()
        self._failed = False
        self._closed = False
        if library is None:
            L = _lib.lib()
        else:
            L = ct.CDLL(str(library))
            L.edt_pair_create.argtypes = [ct.c_int] * 3
            L.edt_pair_create.restype = ct.c_void_p
            L.edt_pair_query.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_void_p, ct.c_void_p]
            L.edt_pair_query.restype = ct.c_int
            L.edt_pair_destroy.argtypes = [ct.c_void_p]
            L.edt_pair_destroy.restype = None
        self._library = L
        create = L.edt_pair_create if capped else L.fe_edt_pair_create
        self._pointer = create(*self.shape)
        if not self._pointer:
            raise RuntimeError('Persistent EDT allocation failed')

    def _check(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError('EDT belongs to another thread')
        if self._closed or self._failed:
            raise RuntimeError('EDT closed or failed')

    def query(self, mask, pitch, *, with_squared=False):
        self._check()
        (validate_large if self._capped else validate_volume)(mask)
        if mask.shape != self.shape:
            raise ValueError('Mask shape differs from prepared shape')
        if isinstance(pitch, (bool, np.bool_)) or not isinstance(pitch, Real) or not np.isfinite(pitch) or \
                not 1e-12 <= pitch <= 1e12:
            raise ValueError('Bounded real pitch required')
        source = np.ascontiguousarray(mask, dtype=np.uint8)
        output = np.empty(self.shape, dtype=np.float32)
        squared = np.empty((2,) + self.shape, dtype=np.int32) if with_squared else None
        status = self._library.edt_pair_query(self._pointer, source.ctypes.data, float(pitch), output.ctypes.data,
                                              squared.ctypes.data if squared is not None else None)
        if status:
            self._failed = status != 1
            raise RuntimeError(f'Persistent EDT status {status}' + _detail(status))
        return (output, squared) if with_squared else output

    def close(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError('EDT belongs to another thread')
        if not self._closed:
            self._library.edt_pair_destroy(self._pointer)
            self._pointer = None
            self._closed = True

    def __enter__(self):
        self._check()
        return self

    def __exit__(self, *args):
        self.close()


# ---------------------------------------------------------------- device-resident
def edt_device(mask, planes=2, out=None, work=None, stream=None):
    """EDT of a uint8 torch tensor (nx,ny,nz) already on the GPU.

    planes=2: returns int32 (2,nx,ny,nz), plane 0 = squared distance of
    non-mask voxels to the mask, plane 1 = of mask voxels to the complement.
    planes=1: int32 (nx,ny,nz), squared distance of nonzero voxels to the
    nearest zero voxel.  `out`/`work` may be preallocated int32 tensors of the
    result's size; the call is asynchronous on `stream`."""
    import torch
    if mask.dtype != torch.uint8 or mask.dim() != 3 or not mask.is_cuda or not mask.is_contiguous():
        raise ValueError('Contiguous 3D uint8 CUDA tensor required')
    shape = tuple(mask.shape)
    full = (planes,) + shape if planes == 2 else shape
    if out is None:
        out = torch.empty(full, dtype=torch.int32, device=mask.device)
    if work is None:
        work = torch.empty(full, dtype=torch.int32, device=mask.device)
    status = _lib.lib().fe_edt_device(_lib.ptr(mask), *shape, planes, _lib.ptr(work), _lib.ptr(out),
                                      _lib.stream_handle(stream))
    if status:
        raise RuntimeError(f'Device EDT status {status}' + _detail(status))
    return out


def signed_from_pair_device(pair, pitch, out=None, stream=None):
    """Signed distance (float32, nx,ny,nz) from a device pair (2,nx,ny,nz)."""
    import torch
    pitch = _pitch(pitch)
    if out is None:
        out = torch.empty(pair.shape[1:], dtype=torch.float32, device=pair.device)
    status = _lib.lib().fe_edt_signed_device(_lib.ptr(pair), out.numel(), pitch, _lib.ptr(out),
                                             _lib.stream_handle(stream))
    if status:
        raise RuntimeError(f'Device signed EDT status {status}')
    return out


class EDTBatch:
    """Exact EDT of many uneven volumes with one launch per axis pass.

    The volumes' masks live back to back in one uint8 device arena (each at a
    256-byte boundary) and their results in one int32 arena (planes of 64-
    element multiples), so `run()` is three kernel launches for the whole
    batch and `capture()` puts them in one graph.  planes=2 gives each volume
    the pair of `edt_device` (plane 0 non-mask to mask, plane 1 mask to
    complement); planes=1 the squared distance of nonzero voxels to zero.
    No reference counterpart: the reference runs one edt_pair_query per
    volume (FE/edt_persistent_pair_v1/edt.cu:59-78)."""

    def __init__(self, shapes, planes=2, device="cuda:0"):
        import torch
        if planes not in (1, 2):
            raise ValueError('planes must be 1 or 2')
        shapes = [tuple(int(x) for x in s) for s in shapes]
        if not shapes or any(len(s) != 3 or min(s) < 1 or max(s) > 4096 for s in shapes):
            raise ValueError('Nonempty list of 3D shapes with axes in [1, 4096] required')
        self.shapes, self.planes = shapes, planes
        sizes = [int(np.prod(s)) for s in shapes]
        self.mask_off = np.zeros(len(shapes), np.int64)
        self.data_off = np.zeros(len(shapes), np.int64)
        self.plane_stride = np.array([-(-n // 64) * 64 for n in sizes], np.int64)
        m = d = 0
        for i, n in enumerate(sizes):
            self.mask_off[i], self.data_off[i] = m, d
            m += -(-n // 256) * 256
            d += planes * int(self.plane_stride[i])
        self.voxels = sum(sizes)
        self.masks = torch.zeros(m, dtype=torch.uint8, device=device)
        self.out = torch.zeros(d, dtype=torch.int32, device=device)
        self.work = torch.empty(d, dtype=torch.int32, device=device)
        dims = np.ascontiguousarray(shapes, dtype=np.int32)
        status = ct.c_int(0)
        L = _lib.lib()
        self._handle = L.fe_edt_batch_create(dims.ctypes.data, self.mask_off.ctypes.data, self.data_off.ctypes.data,
                                             self.plane_stride.ctypes.data, len(shapes), planes, m, d,
                                             ct.byref(status))
        if not self._handle:
            raise (ValueError if status.value == 1 else RuntimeError)(
                f'EDT batch creation status {status.value}' + _detail(status.value))

    def __len__(self):
        return len(self.shapes)

    def mask(self, i):
        """uint8 device view (nx,ny,nz) of volume i's mask (write it before run)."""
        n = int(np.prod(self.shapes[i]))
        o = int(self.mask_off[i])
        return self.masks[o:o + n].view(self.shapes[i])

    def set_masks(self, volumes):
        """Copies host uint8 arrays (or device tensors) into the mask arena."""
        import torch
        for i, v in enumerate(volumes):
            src = v if isinstance(v, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(v, np.uint8))
            if tuple(src.shape) != self.shapes[i]:
                raise ValueError(f'volume {i}: shape {tuple(src.shape)} != {self.shapes[i]}')
            self.mask(i).copy_(src.to(self.masks.device, non_blocking=False))

    def result(self, i):
        """int32 device view of volume i: (2,nx,ny,nz) for planes=2, else (nx,ny,nz)."""
        n = int(np.prod(self.shapes[i]))
        o, s = int(self.data_off[i]), int(self.plane_stride[i])
        if self.planes == 1:
            return self.out[o:o + n].view(self.shapes[i])
        return self.out[o:o + 2 * s].view(2, s)[:, :n].view((2,) + self.shapes[i])

    def run(self, stream=None):
        if not self._handle:
            raise RuntimeError('EDT batch closed')
        status = _lib.lib().fe_edt_batch_run(self._handle, _lib.ptr(self.masks), _lib.ptr(self.work),
                                             _lib.ptr(self.out), _lib.stream_handle(stream))
        if status:
            raise RuntimeError(f'EDT batch status {status}' + _detail(status))

    def capture(self):
        """A CUDA graph whose replay runs the whole batch (one graph launch)."""
        import torch
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.run()
        return graph

    def close(self):
        if getattr(self, '_handle', None):
            _lib.lib().fe_edt_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
