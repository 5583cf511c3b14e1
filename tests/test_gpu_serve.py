"""Continuous-batching server: every request's predictions are bitwise the same when served
alone and when served in a mix with other requests and a concurrent refit."""

import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(kernel, precision, n=3000, seed=7):
    from qml_amd import krr, synthetic

    train, test, size = synthetic.workload("krr_gaussian_20k", n_train=n, n_test=600)
    Z, R, off = train.ragged()
    m = krr.KRR(kernel, sigma=4000.0 if kernel == "gaussian" else 1000.0, lam=1e-6,
                precision=precision)
    m.fit_molecules(Z, R, off, size, train.labels)
    return m, test, size


@pytest.mark.parametrize("kernel,precision", [("gaussian", "fp64"), ("gaussian", "mixed"),
                                              ("laplacian", "fp64")])
def test_served_equals_alone(kernel, precision):
    from qml_amd import serve

    model, test, size = _model(kernel, precision)
    rng = np.random.default_rng(11)
    cuts = [0]
    while cuts[-1] < test.nmol:
        cuts.append(min(test.nmol, cuts[-1] + int(np.exp(rng.uniform(0, np.log(200))))))
    subsets = [test.subset(a, b) for a, b in zip(cuts[:-1], cuts[1:])]
    alone = []
    for sub in subsets:
        Z, R, off = sub.ragged()
        alone.append(model.predict_molecules(Z, R, off, size))
    with serve.KRRServer(model, size, max_molecules=256, max_inflight=3) as srv:
        reqs = []
        for sub in subsets:  # burst: many requests share batches
            Z, R, off = sub.ragged()
            reqs.append(srv.submit(serve.PredictRequest(Z, R, off)))
        for r in reqs:
            r.wait(60)
        assert srv.batches < len(subsets)  # requests were actually batched together
    for r, a in zip(reqs, alone):
        assert r.result.shape == a.shape
        assert np.array_equal(r.result.view(np.int64), a.view(np.int64))


def test_refit_in_mix_and_admission():
    """A refit runs on its own stream while predictions stream through; its model equals a
    standalone fit bit for bit, and the predictions stay bitwise unchanged."""
    from qml_amd import krr, serve, synthetic

    model, test, size = _model("gaussian", "fp64", n=2000)
    train2, _, _ = synthetic.workload("krr_gaussian_20k", n_train=2500, n_test=0)
    Z2, R2, off2 = train2.ragged()
    solo = krr.KRR("gaussian", sigma=4000.0, lam=1e-6)
    solo.fit_molecules(Z2, R2, off2, size, train2.labels)
    alpha_solo = solo.alpha.cpu().numpy()

    Zt, Rt, offt = test.ragged()
    alone = model.predict_molecules(Zt, Rt, offt, size)
    gate = threading.Event()
    probes = []

    def mem_info():
        probes.append(gate.is_set())
        return ((64 << 30) if gate.is_set() else 0, 288 << 30)

    adm = serve.MemoryAdmission(mem_info, reserve_bytes=0)
    with serve.KRRServer(model, size, max_molecules=300, max_inflight=2, admission=adm) as srv:
        refit = srv.submit(serve.RefitRequest(Z2, R2, off2, train2.labels, size=size))
        time.sleep(0.1)
        assert refit.t_start is None  # held: no memory
        gate.set()
        preds = []
        for k in range(20):
            a, b = (k * 37) % test.nmol, min(test.nmol, (k * 37) % test.nmol + 50)
            sub = test.subset(a, b)
            Z, R, off = sub.ragged()
            preds.append((a, b, srv.submit(serve.PredictRequest(Z, R, off))))
        fitted = refit.wait(120)
        for a, b, r in preds:
            r.wait(60)
            assert np.array_equal(r.result.view(np.int64), alone[a:b].view(np.int64))
    assert probes[0] is False and probes[-1] is True
    assert np.array_equal(fitted.alpha.cpu().numpy().view(np.int64), alpha_solo.view(np.int64))


def test_poisson_load_latency():
    from qml_amd import serve

    model, test, size = _model("gaussian", "fp64", n=2000)
    rng = np.random.default_rng(5)
    with serve.KRRServer(model, size, max_molecules=1024, max_inflight=3) as srv:
        reqs = serve.poisson_load(srv, test, rate=400.0, duration=1.0, rng=rng, max_mol=128,
                                  size=size)
        for r in reqs:
            r.wait(60)
    lat = np.array([r.latency for r in reqs])
    assert len(reqs) > 100 and np.all(np.isfinite(lat)) and lat.max() < 5.0
