"""Mesh stage: triangle mesh -> (gmin, shape, surface, solid, signed distance).

`PreparedMeshField` and `surface_raster_and_flood_cuda` keep the constructor
arguments, validation messages, method selection, outputs and error types of
FE/mesh_field_cuda_columns_v1.py:20-98.  Preparation binds the immutable
geometry and grid (on the device, with the column scan's tile lists built
once); every `evaluate` recomputes signing and distances by replaying one
captured HIP graph:

* closed, winding-consistent meshes ("cuda_columns"): +z column hits ->
  composite-key mask -> 6-neighbour surface -> paired EDT -> signed distance;
* any other mesh ("ordered_gwn"): |ordered generalized winding number| > .5 at
  every voxel sample, then the same surface and distances.

The native library only accepts what these wrappers hand it; there is no host
fallback.
"""
import ctypes as ct
import operator
import threading
from numbers import Real

import numpy as np

from . import _lib

# Voxel i of the local grid samples anchor + (i + VOXEL_SAMPLE) * pitch; x and y
# are shifted by pitch * JITTER_FRAC * GOLDEN[:2] to break ties with mesh edges
# (FE/faltkarna_v1_mesh_to_sdf.py:122-131; the same numbers, so samples land
# on the same doubles).
GOLDEN = (0.3819660112501051, 0.6180339887498949, 0.2360679774997897)
JITTER_FRAC = 1.0e-4
VOXEL_SAMPLE = 0.0

HIT_CAPACITY = 64

_STATUS_TEXT = {4: "column query status 4"}


def grid_window(vertices, pitch, lo, global_shape=None, margin_vox_pad=2):
    """(gmin, shape) of the local window in global grid indices
    (FE/faltkarna_v1_mesh_to_sdf.py:153-167)."""
    v = np.asarray(vertices, dtype=np.float64)
    gmin = np.floor((v.min(0) - lo) / pitch).astype(np.int64) - margin_vox_pad
    gmax = np.ceil((v.max(0) - lo) / pitch).astype(np.int64) + margin_vox_pad + 1
    if global_shape is not None:
        gmin = np.maximum(gmin, 0)
        gmax = np.minimum(gmax, np.array(global_shape, dtype=np.int64))
    gmax = np.maximum(gmax, gmin)
    return gmin, tuple(int(x) for x in (gmax - gmin))


def mesh_topology(faces):
    """(watertight, winding_consistent) by vertex index: every undirected edge
    is shared by exactly two faces, and each such pair runs in opposite
    directions (the rule of trimesh's is_watertight / is_winding_consistent
    that FE/mesh_field_cuda_columns_v1.py:45 uses to pick the method)."""
    f = np.asarray(faces, dtype=np.int64)
    edges = f[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2)
    if not len(edges):
        return False, False
    key = np.sort(edges, axis=1)
    order = np.lexsort((key[:, 1], key[:, 0]))
    s = key[order]
    start = np.r_[True, np.any(s[1:] != s[:-1], axis=1)]
    first = np.flatnonzero(start)
    size = np.diff(np.r_[first, len(s)])
    pairs = first[size == 2]
    watertight = 2 * len(pairs) == len(edges)
    a, b = edges[order[pairs]], edges[order[pairs + 1]]
    winding = bool(np.all(a[:, 1] == b[:, 0]))
    return bool(watertight), winding


def _arr(a):
    return a.ctypes.data_as(ct.c_void_p)


class PreparedMeshField:
    """FE/mesh_field_cuda_columns_v1.py:20-98 on the device.  The library
    keyword arguments of the reference are accepted and unused (this
    package's library serves every stage).  The reference's native grid caps
    (<= 512 per axis, <= 1048576 voxels) are lifted to <= 4096 per axis and
    fewer than 2^31 voxels."""

    def __init__(self, vertices, faces, pitch, lo, *, columns_library=None, edt_library=None, column_library=None,
                 global_shape=None, margin_vox_pad=2):
        v = np.array(vertices, dtype=np.float64, copy=True)
        t = np.array(faces, copy=True)
        lo = np.array(lo, dtype=np.float64, copy=True)
        if v.ndim != 2 or v.shape[1] != 3 or not len(v) or not np.isfinite(v).all():
            raise ValueError('finite vertices required')
        if t.ndim != 2 or t.shape[1] != 3 or not len(t) or not np.issubdtype(t.dtype, np.integer):
            raise ValueError('integral faces required')
        if t.min() < 0 or t.max() >= len(v):
            raise ValueError('face index outside vertices')
        if lo.shape != (3,) or not np.isfinite(lo).all():
            raise ValueError('finite three-coordinate origin required')
        if isinstance(pitch, (bool, np.bool_)) or not isinstance(pitch, Real) or not np.isfinite(pitch) or \
                not 1e-12 <= pitch <= 1e12:
            raise ValueError('bounded real pitch required')
        if isinstance(margin_vox_pad, (bool, np.bool_)):
            raise ValueError('integer padding required')
        margin_vox_pad = operator.index(margin_vox_pad)
        if not 0 <= margin_vox_pad <= 512:
            raise ValueError('padding outside contract')
        if global_shape is not None:
            raw = tuple(global_shape)
            if len(raw) != 3 or any(isinstance(x, (bool, np.bool_)) for x in raw):
                raise ValueError('three integer global dimensions required')
            global_shape = tuple(operator.index(x) for x in raw)
            if min(global_shape) < 1 or max(global_shape) > 1000000000:
                raise ValueError('global grid outside index contract')
        relative = (v - lo) / float(pitch)
        if not np.isfinite(relative).all() or np.max(np.abs(relative)) > 1e12:
            raise ValueError('grid coordinates exceed bounded index contract')
        gmin, shape = grid_window(v, float(pitch), lo, global_shape, margin_vox_pad)
        if min(shape) < 1 or max(shape) > 4096 or int(np.prod(shape)) >= 2 ** 31:
            raise ValueError('local grid outside native EDT capacity')
        anchor = lo + gmin.astype(np.float64) * float(pitch)
        tri = v[t.astype(np.int64)]
        cross = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        if np.any(np.sqrt((cross ** 2).sum(axis=1)) / 2 == 0):
            raise ValueError('degenerate triangle')
        watertight, winding = mesh_topology(t)
        self.method = 'cuda_columns' if watertight and winding else 'ordered_gwn'
        self._anchor = anchor
        self._pitch = float(pitch)
        self._gmin = gmin
        self.shape = shape
        self._owner = threading.get_ident()
        self._closed = False
        self._failed = False
        self._handle = ct.c_void_p()
        self._lib = _lib.lib()
        self._gmin.flags.writeable = False
        nx, ny, nz = shape
        if self.method == 'cuda_columns':
            corners = np.ascontiguousarray(tri, dtype=np.float64)
            ij = np.indices(shape[:2]).reshape(2, -1).T
            xy = anchor[:2] + (ij + VOXEL_SAMPLE) * self._pitch
            xy[:, 0] += self._pitch * JITTER_FRAC * GOLDEN[0]
            xy[:, 1] += self._pitch * JITTER_FRAC * GOLDEN[1]
            xy = np.ascontiguousarray(xy)
            zs = np.ascontiguousarray(anchor[2] + (np.arange(nz, dtype=np.float64) + VOXEL_SAMPLE) * self._pitch)
            a = np.ascontiguousarray(anchor)
            status = self._lib.fe_mesh_create(_arr(corners), len(corners), _arr(a), _arr(xy), nx, ny, nz,
                                              self._pitch, _arr(zs), HIT_CAPACITY, ct.byref(self._handle))
        else:
            # Sample points and corners relative to the anchor, computed as the
            # reference does (world grid first, then the subtraction in f64).
            points = anchor + np.indices(shape).reshape(3, -1).T * self._pitch
            points[:, 0] += self._pitch * JITTER_FRAC * GOLDEN[0]
            points[:, 1] += self._pitch * JITTER_FRAC * GOLDEN[1]
            points = np.ascontiguousarray(points - anchor)
            corners = np.ascontiguousarray((v - anchor)[t.astype(np.int64)], dtype=np.float64)
            status = self._lib.fe_gwn_create(_arr(corners), len(corners), _arr(points), nx, ny, nz, self._pitch,
                                             ct.byref(self._handle))
        if status:
            raise RuntimeError(f'column creation status {status}')

    def _check(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError('field belongs to another thread')
        if self._closed:
            raise RuntimeError('field is closed')
        if self._failed:
            raise RuntimeError('field evaluation failed; no retry')

    def evaluate(self):
        """(gmin, shape, surface, solid, distance): bool, bool, float32 arrays of `shape`."""
        self._check()
        solid = np.empty(self.shape, np.uint8)
        surface = np.empty(self.shape, np.uint8)
        distance = np.empty(self.shape, np.float32)
        status = self._lib.fe_mesh_query(self._handle, _arr(solid), _arr(surface), _arr(distance))
        if status:
            self._failed = True
            if status == 1:
                raise ValueError('Both classes must be present')
            if status == 5:
                raise ValueError('query coincides with a vertex')
            raise RuntimeError(_STATUS_TEXT.get(status, f'mesh field status {status}'))
        return self._gmin.copy(), self.shape, surface.view(np.bool_), solid.view(np.bool_), distance

    def evaluate_into(self, solid, surface, distance):
        """evaluate() into caller arrays (bool, bool, float32 of `shape`, C-contiguous,
        writable, not overlapping), as FE/mesh_fused_into_v1.py:43-62; the
        device copies land in them directly."""
        self._check()
        arrays = (solid, surface, distance)
        for array, dtype in zip(arrays, (np.bool_, np.bool_, np.float32)):
            if not isinstance(array, np.ndarray) or array.dtype != dtype or array.shape != tuple(self.shape):
                raise ValueError('Matching output shape and dtype required')
            if not array.flags.c_contiguous or not array.flags.writeable:
                raise ValueError('Contiguous writable outputs required')
        spans = sorted((a.ctypes.data, a.ctypes.data + a.nbytes) for a in arrays)
        if any(left[1] > right[0] for left, right in zip(spans, spans[1:])):
            raise ValueError('Output buffers must not overlap')
        status = self._lib.fe_mesh_query(self._handle, *(_arr(a) for a in arrays))
        if status:
            self._failed = True
            if status == 1:
                raise ValueError('Both classes must be present')
            if status == 5:
                raise ValueError('query coincides with a vertex')
            raise RuntimeError(_STATUS_TEXT.get(status, f'mesh field status {status}'))

    def pin_outputs(self, region):
        """Checks an output region of the service layout (FE/mesh_fused_into_v1.py:8-30)
        and returns False: memory that already exists is not page-locked here,
        so device copies into it are staged by the runtime."""
        self._check()
        size = int(np.prod(self.shape))
        expected = (2 * size + 63) // 64 * 64 + 4 * size
        if len(region) != expected or region.readonly:
            raise ValueError('Writable exact-size region required')
        return False

    def unpin_outputs(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError('Field belongs to another thread')

    def run_device(self, stream=None):
        """Replay the stage on a torch stream (default: the current one)
        without copying anything out (results stay in the handle's device
        buffers, `device_buffers()`)."""
        self._check()
        status = self._lib.fe_mesh_run(self._handle, _lib.stream_handle(stream))
        if status:
            self._failed = True
            raise RuntimeError(f'mesh field status {status}')

    def device_status(self):
        """Status of the last completed replay (0 ok, 1 one class, 4 overflow, 5 corner)."""
        return int(self._lib.fe_mesh_status(self._handle))

    def device_buffers(self):
        """Raw device pointers: solid, surface, distance, pair, counts, keys, signs, and the hit capacity."""
        ptrs = [ct.c_void_p() for _ in range(7)]
        cap = ct.c_int()
        self._lib.fe_mesh_buffers(self._handle, *[ct.byref(p) for p in ptrs], ct.byref(cap))
        return tuple(p.value for p in ptrs) + (cap.value,)

    def close(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError('field belongs to another thread')
        if not self._closed:
            self._closed = True
            status = self._lib.fe_mesh_destroy(ct.byref(self._handle))
            if status:
                raise RuntimeError(f'mesh field close status {status}')

    def __enter__(self):
        self._check()
        return self

    def __exit__(self, *args):
        self.close()


def surface_raster_and_flood_cuda(vertices, faces, pitch, lo, *, columns_library=None, edt_library=None,
                                  column_library=None, global_shape=None, margin_vox_pad=2):
    """One-shot stage including preparation and close
    (FE/mesh_field_cuda_columns_v1.py:91-96)."""
    with PreparedMeshField(vertices, faces, pitch, lo, global_shape=global_shape,
                           margin_vox_pad=margin_vox_pad) as field:
        return field.evaluate()
