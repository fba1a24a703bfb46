"""Host logic of the autotuner cache (no GPU): signatures and the JSON
cache round trip, with a stand-in for the context."""
import json

from mcf_amd import autotune as A
from mcf_amd import workloads as W


class FakeSim:
    def __init__(self, pick=(2, 0), ms=1.5):
        self.pick, self.ms = pick, ms
        self.applied = None
        self.tuned = 0

    def autotune(self, reps):
        self.tuned += 1
        self.applied = self.pick
        return self.pick, self.ms

    def set_tuning(self, shape, order):
        self.applied = (shape, order)


def test_signature_separates_workload_classes():
    a = W.tait_box_3d(16).prm
    s1 = A.signature(a, 16 ** 3, 16.0 ** 3)
    assert s1 == A.signature(a, 16 ** 3 + 10, 16.0 ** 3)          # same class
    assert s1 != A.signature(a, 128 ** 3, 128.0 ** 3)               # scale
    assert s1 != A.signature(a.replace(precision=1), 16 ** 3, 16.0 ** 3)
    assert s1 != A.signature(a.replace(Newtonian=False), 16 ** 3, 16.0 ** 3)
    assert s1 != A.signature(a, 16 ** 3, 16.0 ** 3, nrep=256)


def test_tune_caches_and_reuses(tmp_path):
    path = str(tmp_path / "t.json")
    prm = W.tait_box_3d(16).prm
    sim = FakeSim()
    e1 = A.tune(sim, prm, 4096, 4096.0, cache_path=path)
    assert not e1["cached"] and sim.tuned == 1 and sim.applied == (2, 0)
    data = json.load(open(path))
    assert data[e1["key"]]["shape"] == [2, 0]
    sim2 = FakeSim(pick=(0, 1))
    e2 = A.tune(sim2, prm, 4096, 4096.0, cache_path=path)
    assert e2["cached"] and sim2.tuned == 0 and sim2.applied == (2, 0)
    e3 = A.tune(sim2, prm, 4096, 4096.0, cache_path=path, retune=True)
    assert not e3["cached"] and sim2.tuned == 1 and json.load(open(path))[e3["key"]]["shape"] == [0, 1]


def test_corrupt_cache_is_ignored(tmp_path):
    path = tmp_path / "t.json"
    path.write_text("{not json")
    sim = FakeSim()
    e = A.tune(sim, W.tait_box_3d(16).prm, 4096, 4096.0, cache_path=str(path))
    assert not e["cached"] and sim.tuned == 1
