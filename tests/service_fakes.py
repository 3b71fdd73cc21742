"""CPU stand-in for a prepared mesh field, loaded by the mesh service's worker
process in the CPU tests (the shape of probes/field_engine/field_shared_service_probe.py:19-43)."""
import numpy as np


class FakeStage:
    def __init__(self, vertices, faces, pitch, origin):
        if pitch < 0:
            raise ValueError('Requested preparation failure')
        self.shape = (4, 3, 2)
        self._gmin = np.array(origin, np.int64)
        self.value = float(np.asarray(vertices).sum())
        self.calls = 0

    def evaluate_into(self, solid, surface, distance):
        self.calls += 1
        solid[:] = self.calls % 2 == 1
        surface[:] = not solid.flat[0]
        distance[:] = self.value + self.calls

    def pin_outputs(self, region):
        return False

    def unpin_outputs(self):
        raise AssertionError('the CPU stand-in is never pinned')

    def close(self):
        pass
