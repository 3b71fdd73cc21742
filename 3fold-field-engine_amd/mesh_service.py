"""Mesh-stage service: a spawned worker process owns one prepared mesh and the
GPU; every query recomputes the field into a shared-memory region that the
caller reads through borrowed, read-only views.

Keeps the API, layout and lifetime rules of FE/mesh_field_shared_service_v1.py:
* `layout(shape)`: solid (N bytes), surface (N bytes), padding to a 64-byte
  boundary, distance (N float32) -- one region, :16-19;
* `MeshFieldService(timeout=60, factory=None, **libraries)` with `prepare`,
  `query` -> `FieldLease` (`result` = (gmin, shape, surface, solid, distance)),
  `close`, context manager; one owner thread, one outstanding lease, worker
  errors invalidate the prepared state, a dead or silent worker closes the
  service;
* views keep their mapping alive after `prepare` or `close` (their base owns
  the mapping) and are read-only.

The worker runs this file through runpy (the package directory's name is not
importable), loads the package by path and serves `PreparedMeshField`
(mesh.py), whose `evaluate_into` replays the captured stage and copies the
three arrays straight into the region.  `pin_outputs` returns False: page-
locking memory that already exists is not used on this platform, so the copy
is staged by the runtime.
"""
import multiprocessing as mp
import runpy
import threading
import time
import weakref
from multiprocessing.shared_memory import SharedMemory
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def layout(shape):
    """(voxels, byte offset of the distance array, total bytes) of a result region."""
    size = int(np.prod(shape))
    distance_offset = (2 * size + 63) // 64 * 64
    return size, distance_offset, distance_offset + 4 * size


class _Backing:
    """One shared-memory result region; unlinked when replaced or closed,
    unmapped when the last view of it is gone."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.size, self.distance_offset, total = layout(shape)
        self.shm = SharedMemory(create=True, size=total)
        self.unlinked = False

    def unlink(self):
        if not self.unlinked:
            self.shm.unlink()
            self.unlinked = True

    def __del__(self):
        if hasattr(self, "shm"):
            try:
                self.unlink()
            finally:
                self.shm.close()


class _View:
    """Array-interface owner: the ndarray's base, so the mapping lives as
    long as any view (slices and np.asarray included)."""

    def __init__(self, backing, offset, dtype):
        self.backing = backing
        raw = np.ndarray((backing.shm.size,), np.uint8, buffer=backing.shm.buf)
        self.__array_interface__ = dict(version=3, shape=backing.shape, typestr=np.dtype(dtype).str,
                                        data=(raw.ctypes.data + offset, True))


def _views(backing):
    """(solid, surface, distance) read-only views of a region."""
    return tuple(np.asarray(_View(backing, offset, dtype)) for offset, dtype in
                 ((0, np.bool_), (backing.size, np.bool_), (backing.distance_offset, np.float32)))


def _default_factory():
    import importlib.util
    import sys
    name = "fe_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, HERE / "__init__.py", submodule_search_locations=[str(HERE)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    return sys.modules[name].PreparedMeshField


def _serve(connection, libraries, factory):
    """Worker loop: prepare / attach / query / close over a pipe."""
    field = shared = arrays = None
    pinned = False

    def clear():
        nonlocal field, shared, arrays, pinned
        try:
            if pinned:
                field.unpin_outputs()
        finally:
            pinned = False
            arrays = None
            if shared is not None:
                shared.close()
                shared = None
            if field is not None:
                prior, field = field, None
                prior.close()

    try:
        if factory is None:
            factory = _default_factory()
        while True:
            operation, payload = connection.recv()
            if operation == "close":
                break
            try:
                if operation == "prepare":
                    clear()
                    field = factory(*payload, **libraries)
                    connection.send(("ok", (tuple(field.shape), np.array(field._gmin, copy=True))))
                elif operation == "attach":
                    if field is None or shared is not None:
                        raise RuntimeError("Prepare exactly once before attachment")
                    shared = SharedMemory(name=payload)
                    size, offset, total = layout(field.shape)
                    if shared.size < total:
                        raise ValueError("Shared region size mismatch")
                    arrays = (np.ndarray(field.shape, np.bool_, buffer=shared.buf),
                              np.ndarray(field.shape, np.bool_, buffer=shared.buf, offset=size),
                              np.ndarray(field.shape, np.float32, buffer=shared.buf, offset=offset))
                    pinned = bool(field.pin_outputs(shared.buf[:total]))
                    connection.send(("ok", dict(pinned=pinned)))
                elif operation == "query":
                    if field is None or arrays is None:
                        raise RuntimeError("Prepare and attach before querying")
                    start = time.perf_counter_ns()
                    field.evaluate_into(*arrays)
                    connection.send(("ok", (time.perf_counter_ns() - start) * 1e-6))
                else:
                    raise ValueError("Unknown service operation")
            except Exception as error:
                try:
                    clear()
                except Exception:
                    pass
                connection.send(("error", f"{type(error).__name__}: {error}"))
    except (EOFError, BrokenPipeError):
        pass
    finally:
        try:
            clear()
        finally:
            connection.close()


class FieldLease:
    """A borrowed result: valid until released; the next query reuses the region."""

    def __init__(self, service, token, worker_ms):
        self._service = weakref.ref(service)
        self._token = token
        self._released = False
        solid, surface, distance = _views(service._backing)
        self.result = (service._gmin.copy(), service._backing.shape, surface, solid, distance)
        self.worker_ms = worker_ms

    def release(self):
        service = self._service()
        if service is not None:
            if threading.get_ident() != service._owner:
                raise RuntimeError("Lease belongs to another thread")
            if service._active == self._token:
                service._active = None
        self._released = True

    def __enter__(self):
        if self._released:
            raise RuntimeError("Lease is released")
        return self

    def __exit__(self, *args):
        self.release()


class MeshFieldService:
    """Single owner, one outstanding lease, responses awaited with a timeout."""

    def __init__(self, *, timeout=60.0, factory=None, **libraries):
        if not np.isfinite(timeout) or not 0 < timeout <= 300:
            raise ValueError("Timeout must be in (0, 300] seconds")
        self._owner = threading.get_ident()
        self._timeout = float(timeout)
        self._closed = False
        self._backing = None
        self._active = None
        self._generation = 0
        self._ready = False
        self.metadata = {}
        context = mp.get_context("spawn")
        self._connection, child = context.Pipe()
        self._process = context.Process(
            target=runpy.run_path, args=(str(Path(__file__).resolve()),),
            kwargs=dict(init_globals=dict(_WORKER_ARGS=(child, libraries, factory)), run_name="__fe_mesh_worker__"))
        self._process.start()
        child.close()

    def _check(self, idle=False):
        if threading.get_ident() != self._owner:
            raise RuntimeError("Service belongs to another thread")
        if self._closed:
            raise RuntimeError("Service is closed")
        if idle and self._active is not None:
            raise RuntimeError("Release the outstanding field lease first")

    def _call(self, operation, payload=None):
        self._check()
        try:
            self._connection.send((operation, payload))
            if not self._connection.poll(self._timeout):
                raise TimeoutError("Service response timed out")
            status, value = self._connection.recv()
        except (EOFError, BrokenPipeError, OSError, TimeoutError):
            self.close()
            raise
        if status != "ok":
            self._ready = False
            raise RuntimeError(value)
        return value

    def _drop_backing(self):
        if self._backing is not None:
            self._backing.unlink()
            self._backing = None

    def prepare(self, vertices, faces, pitch, origin):
        """Prepare a mesh in the worker and attach a fresh result region; returns the grid shape."""
        self._check(idle=True)
        self._ready = False
        shape, self._gmin = self._call("prepare", (vertices, faces, pitch, origin))
        self._drop_backing()
        try:
            self._backing = _Backing(shape)
            self.metadata = self._call("attach", self._backing.shm.name)
            self._ready = True
        except Exception:
            self.close()
            raise
        return shape

    def query(self):
        """Recompute the field; the lease's views read the shared region."""
        self._check(idle=True)
        if not self._ready:
            raise RuntimeError("Prepare a field before querying")
        worker_ms = self._call("query")
        self._generation += 1
        self._active = self._generation
        return FieldLease(self, self._active, worker_ms)

    def close(self):
        if threading.get_ident() != self._owner:
            raise RuntimeError("Service belongs to another thread")
        if self._closed:
            return
        try:
            self._connection.send(("close", None))
        except (BrokenPipeError, EOFError, OSError):
            pass
        self._process.join(timeout=10)
        if self._process.is_alive():
            self._process.terminate()
            self._process.join(timeout=5)
        if self._process.is_alive():
            self._process.kill()
            self._process.join(timeout=5)
        if self._process.is_alive():
            raise RuntimeError("Worker did not stop; mapping retained")
        self._closed = True
        self._ready = False
        self._connection.close()
        self._drop_backing()

    def __enter__(self):
        self._check()
        return self

    def __exit__(self, *args):
        self.close()


if __name__ == "__fe_mesh_worker__":
    _serve(*_WORKER_ARGS)  # noqa: F821 (provided by runpy's init_globals)
