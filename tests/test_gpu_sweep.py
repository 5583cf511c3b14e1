"""Batched sweep (many uneven KRR fits in one graph launch) against the oracle."""

import numpy as np
import pytest

import oracle
from qml_amd import synthetic

pytestmark = pytest.mark.gpu

# (n, sigma, lambda, kernel, first pool row): edge sizes around the 64/128/256 blockings
PROBLEMS = [
    (1, 724.0, 1e-8, "laplacian", 0),
    (17, 500.0, 1e-10, "laplacian", 5),
    (64, 724.0, 1e-8, "laplacian", 0),
    (65, 1000.0, 1e-6, "gaussian", 11),
    (127, 1000.0, 1e-6, "gaussian", 3),
    (128, 2000.0, 1e-4, "laplacian", 0),
    (129, 500.0, 1e-10, "laplacian", 7),
    (300, 2000.0, 1e-4, "gaussian", 0),
    (513, 724.0, 1e-8, "laplacian", 100),
    (700, 1000.0, 1e-6, "laplacian", 0),
    (1000, 2000.0, 1e-4, "gaussian", 200),
    (1500, 500.0, 1e-10, "laplacian", 0),
]

_POOL = {}


def _pool():
    if not _POOL:
        from qml_amd.device import rows_to_device

        tr, _, size = synthetic.workload("krr_sweep_64", n_train=1700, n_test=0)
        X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
        _POOL.update(X=X, y=tr.labels.copy(), Xd=rows_to_device(X), d=X.shape[1])
    return _POOL


def _problems(spec):
    from qml_amd.sweep import SweepProblem

    return [SweepProblem(n, s, l, k, x0) for n, s, l, k, x0 in spec]


def _oracle_system(p, X, y):
    kern = oracle.laplacian_kernel if p.kernel == "laplacian" else oracle.gaussian_kernel
    Xp = X[p.x_row0:p.x_row0 + p.n]
    K = kern(Xp, Xp, p.sigma)
    K[np.diag_indices_from(K)] += p.lam
    return K, y[p.x_row0:p.x_row0 + p.n]


def _check_solution(K, y, alpha, alpha_ref):
    # backward error of the solve (independent of conditioning) ...
    r = K @ alpha - y
    berr = np.max(np.abs(r)) / (np.max(np.abs(K).sum(1)) * np.max(np.abs(alpha)) + np.max(np.abs(y)))
    assert berr <= 1e-13 * max(1, len(y)) ** 0.5, berr
    # ... and the forward error against LAPACK within what the condition number allows
    cond = np.linalg.cond(K)
    fwd = np.linalg.norm(alpha - alpha_ref) / np.linalg.norm(alpha_ref)
    assert fwd <= max(1e-12, 64 * len(y) * cond * 2.2e-16), (fwd, cond)


@pytest.mark.parametrize("nb,branches", [(64, 1), (128, 1), (256, 1), (512, 1), (64, 2),
                                         (128, 3), (256, 4), (512, 2)])
def test_batched_matches_oracle(nb, branches):
    from qml_amd.sweep import BatchedFit

    P = _pool()
    probs = _problems(PROBLEMS)
    fit = BatchedFit(probs, P["Xd"], P["d"], P["y"], nb=nb, graph=True, branches=branches).run()
    assert np.all(fit.info() == 0)
    for p, a in zip(probs, fit.alphas()):
        K, y = _oracle_system(p, P["X"], P["y"])
        alpha_ref = oracle.cho_solve(K.copy(), y.copy())
        _check_solution(K, y, a.cpu().numpy(), alpha_ref)
    fit.close()


def test_batched_factor_matches_lapack():
    from scipy.linalg import lapack

    from qml_amd.sweep import BatchedFit

    P = _pool()
    probs = _problems(PROBLEMS)
    fit = BatchedFit(probs, P["Xd"], P["d"], P["y"], nb=128, graph=False, branches=2).run()
    for i in (3, 9, 11):
        p = probs[i]
        K, _ = _oracle_system(p, P["X"], P["y"])
        Lref, info = lapack.dpotrf(K, lower=1, clean=1)
        assert info == 0
        F = fit.factor(i).cpu().numpy()[:, :p.n].T  # rows of the view are columns
        Lg = np.tril(F)
        assert np.linalg.norm(Lg - Lref) / np.linalg.norm(Lref) <= 1e-12
    fit.close()


def test_each_result_same_alone_and_in_batch():
    from qml_amd.sweep import BatchedFit

    P = _pool()
    probs = _problems(PROBLEMS)
    # the batch in four stream groups, each problem alone on one stream
    full = BatchedFit(probs, P["Xd"], P["d"], P["y"], nb=128, graph=True, branches=4).run()
    for i in (0, 4, 8, 11):
        alone = BatchedFit([probs[i]], P["Xd"], P["d"], P["y"], nb=128, graph=True,
                           branches=1).run()
        a = alone.alphas()[0].cpu().numpy()
        b = full.alphas()[i].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), i
        alone.close()
    full.close()


def test_graph_replays_are_identical():
    import torch

    from qml_amd.sweep import BatchedFit

    P = _pool()
    probs = _problems(PROBLEMS[4:])
    fit = BatchedFit(probs, P["Xd"], P["d"], P["y"], nb=64, graph=True, branches=2)
    fit.run_direct()
    ref = fit.alpha.clone()
    for _ in range(3):
        fit.alpha.zero_()
        junk = torch.randn(512, 512, dtype=torch.float64, device=ref.device)
        (junk @ junk).sum().item()  # eager work and a sync between replays
        fit.run()
        assert torch.equal(fit.alpha, ref)
        assert np.all(fit.info() == 0)
    fit.close()


def test_info_matches_lapack_and_isolates_failure():
    from scipy.linalg import lapack

    from qml_amd.sweep import BatchedFit, SweepProblem

    P = _pool()
    good = _problems(PROBLEMS[6:10])
    bad = SweepProblem(400, 2000.0, -0.5, "laplacian", 0)
    fit = BatchedFit(good + [bad], P["Xd"], P["d"], P["y"], nb=128, graph=True,
                     branches=2).run()
    K, _ = _oracle_system(bad, P["X"], P["y"])
    _, info_ref = lapack.dpotrf(K, lower=1)
    info = fit.info()
    assert info_ref > 0 and info[-1] == info_ref and np.all(info[:-1] == 0)
    clean = BatchedFit(good, P["Xd"], P["d"], P["y"], nb=128, graph=True, branches=1).run()
    for a, b in zip(fit.alphas()[:-1], clean.alphas()):
        assert np.array_equal(a.cpu().numpy().view(np.uint64), b.cpu().numpy().view(np.uint64))
    fit.close()
    clean.close()


def test_batched_vs_sequential_and_autotune(tmp_path):
    import torch

    from qml_amd import _lib
    from qml_amd.autotune import Autotuner
    from qml_amd.kernels import pairwise_device
    from qml_amd.sweep import (BatchedFit, SCHEDULE_CANDIDATES, learning_curve_grid,
                               schedule_name, sequential_fit)

    P = _pool()
    probs = learning_curve_grid(count=16, n_min=200, n_max=1700, seed=3)
    tuner = Autotuner(str(tmp_path / "tune.json"))
    fit = BatchedFit(probs, P["Xd"], P["d"], P["y"], nb=None, tuner=tuner).run()
    choice = schedule_name(fit.nb, fit.branches)
    assert choice in SCHEDULE_CANDIDATES and tuner.lookup(list(tuner._read())[0]) == choice
    assert len(tuner._read()[list(tuner._read())[0]]["timings"]) == len(SCHEDULE_CANDIDATES)
    seq = sequential_fit(probs, P["Xd"], P["d"], P["y"])
    torch.cuda.synchronize()
    yd = torch.from_numpy(P["y"]).cuda()
    for p, a, s in zip(probs, fit.alphas(), seq):
        X = P["Xd"][:p.n]
        K = pairwise_device(_lib.PAIR_LAPLACIAN, X, p.n, X, p.n, P["d"], p.sigma, True)
        K.diagonal().add_(p.lam)
        for alpha in (a, s):
            r = K @ alpha - yd[:p.n]
            den = K.abs().sum(1).max() * alpha.abs().max() + yd[:p.n].abs().max()
            assert float(r.abs().max() / den) <= 1e-13 * p.n ** 0.5
    fit.close()
