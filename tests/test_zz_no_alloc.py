"""The allocation-free prediction entry point (qmlh_krr_predict_noalloc) against the one that
takes a workspace (qmlh_krr_predict): bitwise equal results on two calls in a row, the second
smaller, for both kernels; the scratch-size limit is an error, not a fault."""

import numpy as np
import pytest

from qml_amd import synthetic

pytestmark = pytest.mark.gpu


def _fit(kernel, sigma, n=3000):
    from qml_amd import krr

    tr, te, size = synthetic.workload("krr_gaussian_20k", n_train=n, n_test=900)
    m = krr.KRR(kernel, sigma, 1e-6)
    m.fit_molecules(*tr.ragged(), size, tr.labels)
    return m, te, size


def _queries(model, te, size, a, b):
    from qml_amd.device import ragged_to_device
    from qml_amd.kernels import row_sqnorms_device
    from qml_amd.representations import generate_coulomb_matrices_device

    Qd, _ = generate_coulomb_matrices_device(*ragged_to_device(*te.subset(a, b).ragged()), size,
                                             ld=model.Xd.shape[1])
    qn = row_sqnorms_device(Qd, b - a, model.d) if model.kernel == "gaussian" else None
    return Qd, qn


@pytest.mark.parametrize("kernel,sigma", [("gaussian", 4000.0), ("laplacian", 724.0)])
def test_noalloc_equals_workspace_entry_point(kernel, sigma):
    import torch

    from qml_amd import _lib
    from qml_amd.krr import KERNELS

    L = _lib.lib()
    model, te, size = _fit(kernel, sigma)
    h = _lib.stream_handle()
    for a, b in ((0, 700), (700, 830)):  # two calls in a row, the second with fewer queries
        m = b - a
        Qd, qn = _queries(model, te, size, a, b)
        ws = torch.empty(L.qmlh_predict_workspace(m, model.n) // 8 + 1, dtype=torch.float64,
                         device="cuda")
        y_ws = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
        y_na = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
        args = (KERNELS[kernel], _lib.ptr(Qd), m, Qd.shape[1], _lib.ptr(qn), _lib.ptr(model.Xd),
                model.n, model.Xd.shape[1], _lib.ptr(model.norms), _lib.ptr(model.alpha),
                model.d, sigma)
        _lib.check(L.qmlh_krr_predict(*args, _lib.ptr(y_ws), _lib.ptr(ws), ws.numel() * 8, h))
        _lib.check(L.qmlh_krr_predict_noalloc(*args, _lib.ptr(y_na), h))
        torch.cuda.synchronize()
        a_ws, a_na = y_ws.cpu().numpy(), y_na.cpu().numpy()
        assert np.all(np.isfinite(a_na))
        assert np.array_equal(a_ws.view(np.int64), a_na.view(np.int64)), (kernel, a, b)


def test_noalloc_limits_and_empty_training_set():
    import torch

    from qml_amd import _lib

    L = _lib.lib()
    cap = L.qmlh_krr_predict_noalloc_capacity()
    assert cap == 16 << 20
    n = 3000
    chunks = -(-n // 128)
    m_over = cap // 8 // chunks + 1  # one query too many for the scratch
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = _lib.ptr(buf)
    rc = L.qmlh_krr_predict_noalloc(_lib.PAIR_LAPLACIAN, p, m_over, 448, None, p, n, 448, None, p,
                                    435, 724.0, p, _lib.stream_handle())
    assert rc == -1 and b"scratch" in L.qmlh_last_error()  # QMLH_ERR_ARG
    # no training rows: every prediction is zero (written by a kernel, no memset)
    y = torch.full((37,), 5.0, dtype=torch.float64, device="cuda")
    _lib.check(L.qmlh_krr_predict_noalloc(_lib.PAIR_LAPLACIAN, p, 37, 448, None, p, 0, 448, None,
                                          p, 435, 724.0, _lib.ptr(y), _lib.stream_handle()))
    torch.cuda.synchronize()
    assert torch.all(y == 0.0)
