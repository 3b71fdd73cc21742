"""Mesh-field service on the CPU: the result layout, ownership, lease and
mapping-lifetime rules of FE/mesh_field_shared_service_v1.py, with a CPU stand-in
stage in the worker process (the controls of
probes/field_engine/field_shared_service_probe.py), and the output checks of
PreparedMeshField.evaluate_into / pin_outputs without a device."""
import ctypes as ct
import gc
import threading
import weakref
from multiprocessing.shared_memory import SharedMemory
from types import SimpleNamespace

import numpy as np
import pytest

from service_fakes import FakeStage


def test_layout(fe):
    assert fe.mesh_service.layout((4, 3, 2)) == (24, 64, 64 + 96)
    assert fe.mesh_service.layout((8, 8, 8)) == (512, 1024, 1024 + 2048)
    assert fe.mesh_service.layout((1, 1, 1)) == (1, 64, 68)


def test_timeouts_refused(fe):
    for bad in (0, -1, float('nan'), float('inf'), 301):
        with pytest.raises(ValueError):
            fe.MeshFieldService(timeout=bad, factory=FakeStage)


def expect(errors, call):
    try:
        call()
    except (RuntimeError, ValueError, FileNotFoundError, TimeoutError, EOFError, BrokenPipeError):
        return True
    errors.append(call)
    return False


def test_service_lifecycle(fe):
    errors = []
    service = fe.MeshFieldService(factory=FakeStage)
    try:
        expect(errors, service.query)
        assert service.prepare([2], [], 1, [1, 2, 3]) == (4, 3, 2)
        assert service.metadata['pinned'] is False
        lease = service.query()
        gmin, shape, surface, solid, distance = lease.result
        assert gmin.tolist() == [1, 2, 3] and shape == (4, 3, 2)
        assert solid.all() and not surface.any() and np.all(distance == 3)
        assert not any(a.flags.writeable for a in (solid, surface, distance))
        expect(errors, lambda: distance.setflags(write=True))
        expect(errors, service.query)
        expect(errors, lambda: service.prepare([8], [], 1, [0, 0, 0]))
        for call in (service.query, lease.release, service.close):
            t = threading.Thread(target=lambda c=call: expect(errors, c))
            t.start()
            t.join()
        expect(errors, service.query)
        own = distance.copy()
        lease.release()
        lease.release()
        expect(errors, lease.__enter__)
        with service.query() as second:
            assert np.all(second.result[4] == 4) and not second.result[3].any()
            assert np.all(distance == 4) and np.all(own == 3)  # borrowed region reused
        old_backing = weakref.ref(service._backing)
        old_name = service._backing.shm.name
        retained = np.asarray(distance)[1:].view(np.ndarray)
        del lease, second, surface, solid, distance
        service.prepare([8], [], 1, [0, 0, 0])
        gc.collect()
        assert old_backing() is not None and np.all(retained == 4)  # the view keeps the mapping
        expect(errors, lambda: SharedMemory(name=old_name))  # but the name is unlinked
        del retained
        gc.collect()
        assert old_backing() is None
        with service.query() as replacement:
            assert np.all(replacement.result[4] == 9)
        expect(errors, lambda: service.prepare([0], [], -1, [0, 0, 0]))
        expect(errors, service.query)  # nothing stale after a failed prepare
        service.prepare([11], [], 1, [0, 0, 0])
        live = service.query()
        retained = np.asarray(live.result[4])
        backing = weakref.ref(service._backing)
        service.close()
        service.close()
        assert not service._process.is_alive()
        assert np.all(retained == 12)  # views stay valid after close
        live.release()
        del live, retained
        gc.collect()
        assert backing() is None
        expect(errors, service.query)
        expect(errors, lambda: service.prepare([], [], 1, [0, 0, 0]))
    finally:
        service.close()
    assert not errors, errors


def test_worker_death_closes_the_service(fe):
    with fe.MeshFieldService(factory=FakeStage) as killed:
        killed.prepare([1], [], 1, [0, 0, 0])
        killed._process.terminate()
        killed._process.join(5)
        with pytest.raises((RuntimeError, EOFError, BrokenPipeError, OSError)):
            killed.query()
        assert killed._closed


def test_evaluate_into_checks_outputs_without_a_device(fe):
    stage = fe.PreparedMeshField.__new__(fe.PreparedMeshField)
    stage.shape = (2, 2, 2)
    stage._owner = threading.get_ident()
    stage._closed = stage._failed = False
    stage._handle = ct.c_void_p()
    calls = []

    def fake_query(handle, solid_ptr, surface_ptr, distance_ptr):
        calls.append(True)
        np.ctypeslib.as_array((ct.c_uint8 * 8).from_address(solid_ptr.value))[:] = 1
        np.ctypeslib.as_array((ct.c_uint8 * 8).from_address(surface_ptr.value))[:] = 0
        np.ctypeslib.as_array((ct.c_float * 8).from_address(distance_ptr.value))[:] = 2.5
        return 0

    stage._lib = SimpleNamespace(fe_mesh_query=fake_query)
    a, b = np.empty(stage.shape, bool), np.empty(stage.shape, bool)
    d = np.empty(stage.shape, np.float32)
    for bad in ((a, a, d), (a, b, d.astype(np.float64)), (a, b, d.reshape(4, 2)), (a, b, d[:, :, ::-1])):
        with pytest.raises(ValueError):
            stage.evaluate_into(*bad)
    d.flags.writeable = False
    with pytest.raises(ValueError):
        stage.evaluate_into(a, b, d)
    d.flags.writeable = True
    with pytest.raises(ValueError):
        stage.pin_outputs(memoryview(bytearray(1)))
    assert stage.pin_outputs(memoryview(bytearray(fe.mesh_service.layout(stage.shape)[2]))) is False
    assert not calls
    stage.evaluate_into(a, b, d)
    assert a.all() and not b.any() and np.all(d == 2.5)
    stage._lib = SimpleNamespace(fe_mesh_query=lambda *args: 4)
    with pytest.raises(RuntimeError):
        stage.evaluate_into(a, b, d)
    assert stage._failed
    with pytest.raises(RuntimeError):
        stage.evaluate_into(a, b, d)
