"""Pins of the FP64 oracle of the learned Poisson initial guess network: the analytic
gradient against central finite differences of the loss, and the message passing against
a dense restatement."""
import numpy as np

from helpers import small_hex


def _graph(oracle):
    m = small_hex(n=(4, 3, 3), L=(1.0, 1.0, 1.0), cyclic=(False, True, True))
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, {"xmin": ("fixedValue", 0.0)})
    n = rm.n_cells
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n); e[j] = 1
        A[:, j] = P.apply(e)
    nbrs = [[j for j in range(n) if j != i and A[i, j] != 0] for i in range(n)]
    K = max(len(v) for v in nbrs)
    nb = -np.ones((n, K), np.int32)
    for i, v in enumerate(nbrs):
        nb[i, :len(v)] = v
    return nb


def test_gnn_gradient_finite_differences(oracle):
    rng = np.random.default_rng(0)
    nb = _graph(oracle)
    n = nb.shape[0]
    H, L, fin = 8, 3, 3
    npar = oracle.gnn_n_params(fin, H, L)
    th = rng.normal(size=npar) * 0.4
    x = rng.normal(size=(n, fin)); t = rng.normal(size=n)
    loss, g, _ = oracle.gnn_loss_grad(nb, x, t, th, H, L)
    eps = 1e-6
    idx = rng.choice(npar, 60, replace=False)
    for k in idx:
        tp = th.copy(); tp[k] += eps
        tm = th.copy(); tm[k] -= eps
        fd = (oracle.gnn_loss_grad(nb, x, t, tp, H, L, False)[0] - oracle.gnn_loss_grad(nb, x, t, tm, H, L, False)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gnn_forward_dense_restatement(oracle):
    rng = np.random.default_rng(1)
    nb = _graph(oracle)
    n = nb.shape[0]
    H, L, fin = 6, 3, 3
    npar = oracle.gnn_n_params(fin, H, L)
    th = rng.normal(size=npar) * 0.5
    x = rng.normal(size=(n, fin))
    _, _, out = oracle.gnn_loss_grad(nb, x, np.zeros(n), th, H, L, False)
    # dense numpy restatement
    o = 0
    We = th[o:o + fin * H].reshape(fin, H); o += fin * H
    be = th[o:o + H]; o += H
    h = np.maximum(x @ We + be, 0)
    Mean = np.zeros((n, n))
    for i in range(n):
        js = nb[i][nb[i] >= 0]
        Mean[i, js] = 1.0 / len(js)
    for _ in range(L):
        W = th[o:o + 2 * H * H].reshape(2 * H, H); o += 2 * H * H
        b = th[o:o + H]; o += H
        z = np.concatenate([h, Mean @ h], axis=1) @ W + b
        h = h + np.maximum(z, 0)
    wd = th[o:o + H]; bd = th[o + H]
    assert np.allclose(out, h @ wd + bd, rtol=1e-12, atol=1e-12)
