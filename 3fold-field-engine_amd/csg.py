"""Tiled min/max CSG over leaf-box packed fields, and the bounded tile reader.

Drop-in entry points (same validation, errors and results):
  evaluate(fields, instructions)     FE/field_csg_tiled_v1.py:31-99 and
                                     FE/field_csg_batch_v1.py:18-79 (library keywords accepted)
  compose(left, right, operation=, distance=)   FE/field_csg_blocks_v1.py:12-43
  gather_tile(field, lower, upper)   FE/field_csg_tiled_v1.py:14-28
  NativeTileReader(library).gather   FE/csg_tile_native_v1.py:9-35
The whole instruction program runs on the device (fe_csg_device) and the
result is packed on the device; the operands' tables never leave HBM.
"""
import ctypes as ct
import operator
import sys
from numbers import Real

import numpy as np

from . import _lib
from .blocks import NODE, AdaptiveSampleBlocks

INSTR = np.dtype([("op", "<i4"), ("left", "<i4"), ("right", "<i4"), ("pad", "<i4"), ("offset", "<f8")])
OPS = {"union": 0, "difference": 1, "offset": 2}


def _torch():
    import torch
    return torch


def _program(fields, instructions):
    """Python-side validation of FE/field_csg_tiled_v1.py:32-67; returns the encoded program."""
    fields, instructions = tuple(fields), tuple(instructions)
    if not 1 <= len(fields) <= 256 or not 1 <= len(instructions) <= 1024:
        raise ValueError('Bounded nonempty inputs and instructions required')
    if any(not isinstance(f, AdaptiveSampleBlocks) for f in fields):
        raise ValueError('Adaptive block operands required')
    first = fields[0]
    if any(f.shape != first.shape or f.pitch != first.pitch or not np.array_equal(f.origin, first.origin)
           for f in fields):
        raise ValueError('Operands must share the exact sampling grid')
    program = []
    for number, instruction in enumerate(instructions):
        if not isinstance(instruction, (tuple, list)) or len(instruction) != 3:
            raise ValueError('Three-item instruction required')
        operation, left, right = instruction
        if operation not in OPS:
            raise ValueError('Unknown operation')

        def reference(value):
            if isinstance(value, (bool, np.bool_)):
                raise ValueError('Integer reference required')
            try:
                index = operator.index(value)
            except TypeError:
                raise ValueError('Integer reference required') from None
            if not 0 <= index < len(fields) + number:
                raise ValueError('Reference must name an input or earlier instruction')
            return index
        left = reference(left)
        if operation == 'offset':
            if isinstance(right, (bool, np.bool_)) or not isinstance(right, Real) or not np.isfinite(right):
                raise ValueError('Finite real offset required')
            right = float(right)
        else:
            right = reference(right)
        program.append((operation, left, right))
    enc = np.zeros(len(program), INSTR)
    for i, (op, left, right) in enumerate(program):
        enc[i]["op"] = OPS[op]
        enc[i]["left"] = left
        enc[i]["offset" if op == "offset" else "right"] = right
    return fields, enc


_WORKSPACE = {}  # device -> uint8 tensor, grown as programs need


def _first_bad(field, shape):
    """fe_csg_check of a field's device table for a grid of `shape`, run once
    per (table, shape) and cached on the field (uploading new host arrays
    invalidates it)."""
    d_nodes, d_payload = field._device()
    key = (field._dev_key, tuple(shape))
    cached = getattr(field, "_csg_check", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    sh = np.array(shape, np.int32)
    out = ct.c_int32(0)
    code = _lib.lib().fe_csg_check(d_nodes.data_ptr(), d_nodes.shape[0], d_payload.data_ptr(), d_payload.numel(),
                                   sh.ctypes.data, ct.byref(out), _lib.stream_handle(None))
    if code == 2:
        raise RuntimeError("Native CSG operand check failed: " + _lib.lib().fe_last_error().decode())
    value = out.value if code == 0 else 0
    field._csg_check = (key, value)
    return value


def evaluate_dense(fields, instructions):
    """Runs the program on the device; returns the dense device float32 result.
    Operands are checked once (fe_csg_check, cached per field); the program
    then runs in a cached device workspace (fe_csg_run)."""
    torch = _torch()
    fields, enc = _program(fields, instructions)
    if sys.byteorder != 'little':
        raise ValueError('Little-endian host required')
    shape = np.array(fields[0].shape, np.int32)
    bad = np.array([_first_bad(f, fields[0].shape) for f in fields], np.int32)
    dev = [f._device() for f in fields]
    n = len(fields)
    nodes = (ct.c_void_p * n)(*[d[0].data_ptr() for d in dev])
    pays = (ct.c_void_p * n)(*[d[1].data_ptr() for d in dev])
    counts = np.array([d[0].shape[0] for d in dev], np.int64)
    lengths = np.array([d[1].numel() for d in dev], np.int64)
    L = _lib.lib()
    need = L.fe_csg_workspace_bytes(n, shape.ctypes.data, enc.ctypes.data, len(enc))
    if need < 0:
        raise ValueError('Native CSG validation failed: program')
    device = dev[0][0].device
    ws = _WORKSPACE.get(device)
    if ws is None or ws.numel() < need:
        ws = _WORKSPACE[device] = torch.empty(max(need, 1), dtype=torch.uint8, device=device)
    out = torch.empty(fields[0].shape, dtype=torch.float32, device=device)
    code = L.fe_csg_run(nodes, counts.ctypes.data, pays, lengths.ctypes.data, n, shape.ctypes.data, enc.ctypes.data,
                        len(enc), bad.ctypes.data, _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle(None))
    if code == 4:
        raise ValueError('Operation exceeds finite float32 range')
    if code == 2:
        raise RuntimeError('Native CSG failed: ' + L.fe_last_error().decode())
    if code:
        raise ValueError(f'Native CSG validation failed: {code}')
    return fields[0], out


def evaluate(fields, instructions, *, reader=None, batch_library=None, pack_library=None, leaf_samples=512):
    """Ordered CSG program over aligned packed operands -> packed result."""
    first, out = evaluate_dense(fields, instructions)
    return AdaptiveSampleBlocks(out, first.origin, first.pitch, leaf_samples=leaf_samples)


def compose(left, right=None, *, operation='union', distance=0.):
    """FE/field_csg_blocks_v1.py:12-43 as a one-instruction program."""
    if not isinstance(left, AdaptiveSampleBlocks):
        raise ValueError('Adaptive block operand required')
    if operation not in OPS:
        raise ValueError('Unknown field operation')
    if operation == 'offset':
        if right is not None or isinstance(distance, (bool, np.bool_)) or not isinstance(distance, Real) or \
                not np.isfinite(distance):
            raise ValueError('Finite real offset and no second operand required')
        return evaluate([left], [('offset', 0, float(distance))])
    if not isinstance(distance, Real) or isinstance(distance, (bool, np.bool_)) or distance != 0:
        raise ValueError('Offset is only valid for the offset operation')
    if not isinstance(right, AdaptiveSampleBlocks):
        raise ValueError('Second adaptive block operand required')
    if left.shape != right.shape or left.pitch != right.pitch or not np.array_equal(left.origin, right.origin):
        raise ValueError('CSG operands must share the exact sampling grid')
    return evaluate([left, right], [(operation, 0, 1)])


class NativeTileReader:
    """FE/csg_tile_native_v1.py:9-35: bounded exact tile copies (csg_gather)."""

    def __init__(self, library=None):
        if sys.byteorder != 'little':
            raise ValueError('Little-endian host required')
        L = _lib.lib() if library is None else ct.CDLL(str(library))
        self._gather = L.csg_gather
        self._gather.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                 ct.c_void_p]
        self._gather.restype = ct.c_int

    def gather(self, field, lower, upper):
        if not isinstance(field, AdaptiveSampleBlocks):
            raise ValueError('Adaptive block field required')
        if field.nodes.dtype != NODE or field.nodes.ndim != 1 or not field.nodes.flags.c_contiguous or \
                field.payload.dtype != np.float32 or field.payload.ndim != 1 or not field.payload.flags.c_contiguous:
            raise ValueError('Contiguous native node and sample arrays required')
        lo, hi = np.asarray(lower), np.asarray(upper)
        if lo.shape != (3,) or hi.shape != (3,) or not np.issubdtype(lo.dtype, np.integer) or \
                not np.issubdtype(hi.dtype, np.integer):
            raise ValueError('Integral tile coordinates required')
        if np.any(lo < 0) or np.any(hi <= lo) or np.any(hi > np.array(field.shape)) or np.prod(hi - lo) > 8192:
            raise ValueError('Tile outside field or bounded capacity')
        lo, hi = np.ascontiguousarray(lo, dtype=np.int32), np.ascontiguousarray(hi, dtype=np.int32)
        result = np.empty(tuple(hi - lo), dtype=np.float32)
        code = self._gather(field.nodes.ctypes.data, len(field.nodes), field.payload.ctypes.data, len(field.payload),
                            lo.ctypes.data, hi.ctypes.data, result.ctypes.data)
        if code:
            raise ValueError(f'Native tile validation failed: {code}')
        return result


def gather_tile(field, lower, upper):
    """FE/field_csg_tiled_v1.py:14-28: an in-bounds tile of a packed field."""
    return NativeTileReader().gather(field, np.asarray(lower), np.asarray(upper))
