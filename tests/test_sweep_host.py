"""Host side of the batched sweep: problem validation in the C ABI (no GPU work happens before
validation), workspace sizing, the problem grid and the autotuner cache."""

import ctypes
import json
import threading

import pytest

from qml_amd import _lib, autotune, sweep


def _probs(spec):
    arr = (_lib.Problem * len(spec))()
    off = 0
    for i, (n, x0, kind) in enumerate(spec):
        arr[i].n, arr[i].x_row0, arr[i].y_offset, arr[i].alpha_offset = n, x0, x0, off
        arr[i].sigma, arr[i].lambda_, arr[i].kind = 724.0, 1e-8, kind
        off += n
    return arr


def test_workspace_size_and_validation():
    L = _lib.load_library()
    good = _probs([(500, 0, _lib.PAIR_LAPLACIAN), (1200, 10, _lib.PAIR_GAUSSIAN)])
    for nb in (64, 128, 256, 512):
        ws = L.qmlh_krr_batched_workspace(good, 2, nb, 1)
        # at least the two factors (ld rounded to 16) and the solve vectors
        assert ws >= 8 * (512 * 500 + 1200 * 1200)
        # two groups: the same factors, job tables split between two launch sequences
        ws2 = L.qmlh_krr_batched_workspace(good, 2, nb, 2)
        assert ws2 >= ws and ws2 - ws < 1 << 20
    assert L.qmlh_krr_batched_workspace(good, 2, 96, 1) == 0  # nb must be 64/128/256/512
    for branches in (0, 5):
        assert L.qmlh_krr_batched_workspace(good, 2, 64, branches) == 0
        assert b"branches" in L.qmlh_last_error()
    bad = _probs([(0, 0, _lib.PAIR_LAPLACIAN)])
    assert L.qmlh_krr_batched_workspace(bad, 1, 64, 1) == 0
    overlap = _probs([(100, 0, _lib.PAIR_LAPLACIAN), (100, 0, _lib.PAIR_LAPLACIAN)])
    overlap[1].alpha_offset = 50
    assert L.qmlh_krr_batched_workspace(overlap, 2, 64, 1) == 0
    assert b"overlap" in L.qmlh_last_error()
    wrong_kind = _probs([(100, 0, _lib.PAIR_MANHATTAN)])
    assert L.qmlh_krr_batched_workspace(wrong_kind, 1, 64, 1) == 0


def test_create_rejects_out_of_range_problems():
    L = _lib.load_library()
    probs = _probs([(300, 100, _lib.PAIR_LAPLACIAN)])
    plan = ctypes.c_void_p()
    fake = 1 << 20  # never dereferenced: validation fails first
    # pool of 350 rows: rows 100..399 run past it
    rc = L.qmlh_krr_batched_create(probs, 1, 64, 1, fake, 350, 280, 276, None, fake, 1000, fake,
                                   1000, fake, fake, 1 << 30, ctypes.byref(plan))
    assert rc != 0 and b"pool" in L.qmlh_last_error() and not plan.value
    rc = L.qmlh_krr_batched_create(probs, 1, 64, 1, fake, 1000, 280, 276, None, fake, 200, fake,
                                   1000, fake, fake, 1 << 30, ctypes.byref(plan))
    assert rc != 0 and b"labels" in L.qmlh_last_error()
    rc = L.qmlh_krr_batched_create(probs, 1, 64, 1, fake, 1000, 280, 276, None, fake, 1000, fake,
                                   1000, fake, fake, 16, ctypes.byref(plan))
    assert rc != 0 and b"workspace" in L.qmlh_last_error()
    g = _probs([(300, 0, _lib.PAIR_GAUSSIAN)])
    rc = L.qmlh_krr_batched_create(g, 1, 64, 1, fake, 1000, 280, 276, None, fake, 1000, fake,
                                   1000, fake, fake, 1 << 30, ctypes.byref(plan))
    assert rc != 0 and b"norms" in L.qmlh_last_error()


def test_learning_curve_grid():
    probs = sweep.learning_curve_grid()
    assert len(probs) == 64
    assert all(500 <= p.n <= 8000 for p in probs)
    pairs = {(p.sigma, p.lam) for p in probs}
    assert len(pairs) == 16  # every (sigma, lambda) grid point appears
    again = sweep.learning_curve_grid()
    assert [p.n for p in probs] == [p.n for p in again]  # seeded


def test_autotuner_cache(tmp_path):
    path = str(tmp_path / "tune.json")
    tuner = autotune.Autotuner(path)
    calls = []

    def measure(c):
        calls.append(c)
        return {64: 3.0, 128: 1.0, 256: 2.0}[c]

    key = autotune.make_key("krr_batched", nmax=8192, count=64)
    assert tuner.tune(key, (64, 128, 256), measure) == 128
    assert calls == [64, 128, 256]
    assert tuner.tune(key, (64, 128, 256), measure) == 128  # cached: not measured again
    assert calls == [64, 128, 256]
    data = json.load(open(path))
    assert data[key]["choice"] == 128 and set(data[key]["timings"]) == {"64", "128", "256"}
    # a cached choice outside the candidate list is ignored and re-tuned
    assert tuner.tune(key, (64, 256), measure) == 256


def test_autotuner_concurrent_writers(tmp_path):
    path = str(tmp_path / "tune.json")

    def work(i):
        autotune.Autotuner(path).tune(f"op|i={i}", (1, 2), lambda c: float(c))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    data = json.load(open(path))
    assert len(data) == 8 and all(v["choice"] == 1 for v in data.values())


def test_tuning_key_buckets():
    a = sweep.tuning_key([sweep.SweepProblem(5000, 724.0, 1e-8)] * 60)
    b = sweep.tuning_key([sweep.SweepProblem(7000, 724.0, 1e-8)] * 64)
    c = sweep.tuning_key([sweep.SweepProblem(9000, 724.0, 1e-8)] * 64)
    assert a == b != c
