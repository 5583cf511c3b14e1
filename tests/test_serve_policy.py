"""Host-side serving policies: batch packing and memory-budget admission of refits."""

import collections

import numpy as np

from qml_amd import serve


class _Req:
    def __init__(self, nmol, natoms):
        self.nmol, self.natoms = nmol, natoms


def test_plan_batch_fifo_and_capacity():
    q = collections.deque(_Req(n, 20 * n) for n in (100, 300, 50, 700, 10))
    b1 = serve.plan_batch(q, 512, 512 * 29)
    assert [r.nmol for r in b1] == [100, 300, 50]  # 700 would overflow; FIFO keeps order
    b2 = serve.plan_batch(q, 512, 512 * 29)
    assert [r.nmol for r in b2] == [700]  # larger than a slot alone: admitted alone
    b3 = serve.plan_batch(q, 512, 512 * 29)
    assert [r.nmol for r in b3] == [10] and not q
    assert serve.plan_batch(q, 512, 512 * 29) == []


def test_plan_batch_atom_cap():
    q = collections.deque(_Req(10, 280) for _ in range(5))
    b = serve.plan_batch(q, 1000, 600)
    assert len(b) == 2 and len(q) == 3


def test_memory_admission_budget():
    free = [10 << 30]
    adm = serve.MemoryAdmission(lambda: (free[0], 288 << 30), reserve_bytes=1 << 30)
    assert adm.admit(9 << 30)
    assert not adm.admit((9 << 30) + 1)
    free[0] = 40 << 30
    assert adm.admit(30 << 30)


def test_refit_bytes_dominated_by_kernel_matrix():
    d = 435
    for prec, per in (("fp64", 8), ("mixed", 12)):
        b = serve.refit_bytes(50000, d, prec)
        assert per * 50000 ** 2 <= b <= per * 50000 ** 2 * 1.02
    assert serve.refit_bytes(1000, d, "mixed") > serve.refit_bytes(1000, d, "fp64")


def test_refit_worker_waits_for_memory():
    """A refit whose footprint exceeds the free budget is held until memory frees up."""
    free = [1 << 30]
    adm = serve.MemoryAdmission(lambda: (free[0], 288 << 30), reserve_bytes=0)
    need = serve.refit_bytes(20000, 435, "fp64")
    decisions = [adm.admit(need)]
    free[0] = need
    decisions.append(adm.admit(need))
    assert decisions == [False, True]


def test_request_ragged_views():
    Z = np.array([1.0, 8.0, 1.0, 6.0])
    R = np.arange(12, dtype=np.float64).reshape(4, 3)
    r = serve.PredictRequest(Z, R, np.array([0, 3, 4]))
    assert r.nmol == 2 and r.natoms == 4


def test_should_launch_coalesces_only_under_load():
    # nothing on the GPU: any queued request goes at once
    assert serve.should_launch(False, 1, 3, 0.0, 128, 2e-4)
    # batches in flight: wait for min_batch molecules or the window
    assert not serve.should_launch(True, 2, 40, 1e-5, 128, 2e-4)
    assert serve.should_launch(True, 5, 200, 1e-5, 128, 2e-4)
    assert serve.should_launch(True, 1, 3, 3e-4, 128, 2e-4)
    # an empty queue never launches; an empty request still leaves after the window
    assert not serve.should_launch(False, 0, 0, 1.0, 128, 2e-4)
    assert serve.should_launch(True, 1, 0, 1e-3, 128, 2e-4)
