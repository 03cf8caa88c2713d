"""Learned Poisson initial guess on the GPU: MFMA BF16 forward and backward kernels against
the FP64 oracle, training on solver-recorded snapshots, and the gain over the warm start."""
import sys

import numpy as np
import pytest

from helpers import EPS0, ME, QE

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plasma_cloud(n=(12, 10, 8), ppc=40, seed=1):
    """3-D electron plasma over a neutralising background between grounded x walls,
    cyclic in y and z, ElectroStatic field solve (a small version of the 1M-cell case)."""
    from picfoam_amd import meshgen, picmi
    L = 1e-3
    mesh = meshgen.hex_box(*n, L, L * n[1] / n[0], L * n[2] / n[0], cyclic=(False, True, True))
    ncell = mesh["n_cells"]
    dens = 1e16
    w = dens * L ** 3 * (n[1] / n[0]) * (n[2] / n[0]) / (ncell * ppc)
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": ncell * ppc}]
    models = {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "poissonTolerance": 1e-8, "seed": seed}
    bc = {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 0.0)}
    c = picmi.Cloud(mesh, species, models, bc)
    c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ncell, dens * QE))
    c.init_maxwellian(0, ppc, 2.0 * 11604.5, seed=seed)
    wp = np.sqrt(dens * QE * QE / (EPS0 * ME))
    return c, 0.1 / wp


def _graph_from_cloud(c):
    """neighbour lists of the Laplacian couplings, read through the operator itself"""
    from oracle import picref
    rm = picref.RefMesh(c.mesh.d)
    P = picref.RefPoisson(rm, {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 0.0)})
    n = rm.n_cells
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n); e[j] = 1.0
        A[:, j] = P.apply(e)
    nbrs = [[j for j in range(n) if j != i and A[i, j] != 0] for i in range(n)]
    K = max(len(v) for v in nbrs)
    nb = -np.ones((n, K), np.int32)
    for i, v in enumerate(nbrs):
        nb[i, :len(v)] = v
    return nb, A


def test_gnn_gradient_matches_oracle(oracle):
    say("gnn gradient vs oracle")
    c, dt = _plasma_cloud(n=(8, 6, 4), ppc=20)
    c.gnn_init(layers=3, width=64, seed=3)
    c.gnn_record(4)
    for _ in range(3):
        c.step(dt)
    assert c.gnn_snapshot_count() == 3
    x, t, norm = c.gnn_snapshot(2)
    theta = c.gnn_params().astype(np.float64)
    # move the parameters off the initialisation so every layer is active
    rng = np.random.default_rng(0)
    theta += rng.normal(size=theta.size) * 0.05
    c.gnn_params(theta.astype(np.float32))
    theta = c.gnn_params().astype(np.float64)
    g_gpu, loss_gpu = c.gnn_grad(2)
    nb, A = _graph_from_cloud(c)
    out, _ = oracle.gnn_forward_backward(nb, x.astype(np.float64), theta, 64, 3)
    e = out - t
    loss_ref = e @ A @ e / norm
    dout = 2 * (A @ e) / norm
    _, g_ref = oracle.gnn_forward_backward(nb, x.astype(np.float64), theta, 64, 3, dout)
    cos = g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref))
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    say(f"loss gpu {loss_gpu:.6f} oracle {loss_ref:.6f}; gradient cosine {cos:.5f}, rel err {rel:.3e}")
    # BF16 activations and weights (8-bit mantissa) bound the agreement
    assert abs(loss_gpu - loss_ref) / loss_ref < 2e-2
    assert cos > 0.995 and rel < 0.1


def test_gnn_training_and_gain():
    say("gnn training")
    c, dt = _plasma_cloud()
    c.gnn_init(layers=3, width=64, seed=5)
    c.gnn_record(64)
    iters_plain = []
    for k in range(80):
        st = c.step(dt)
        if k >= 16:
            iters_plain.append(st.poisson.iterations)
    assert c.gnn_snapshot_count() == 64
    l0 = np.mean([c.gnn_eval(i) for i in range(8)])
    first, last = c.gnn_train(600, lr=2e-3, seed=2)
    l1 = np.mean([c.gnn_eval(i) for i in range(8)])
    say(f"A-norm loss of the guess relative to the warm start: before {l0:.4f} after {l1:.4f} "
        f"(train first {first:.4f} last {last:.4f})")
    assert l1 < l0 and l1 < 1.0
    # use the network as the initial guess on the following steps
    c.gnn_record(0)
    c.gnn_use(True)
    iters_gnn = [c.step(dt).poisson.iterations for _ in range(40)]
    say(f"PCG iterations to 1e-8: warm start {np.mean(iters_plain):.2f}, learned guess {np.mean(iters_gnn):.2f}")
    assert np.mean(iters_gnn) <= np.mean(iters_plain) + 0.5


def test_mfma_layer_matches_bf16_matmul(oracle):
    """The MFMA fragment layout: a one-layer network with random weights on random
    features equals the dense restatement with BF16-rounded operands."""
    say("mfma layout")
    c, dt = _plasma_cloud(n=(8, 6, 4), ppc=10)
    c.gnn_init(layers=1, width=64, seed=9)
    c.gnn_record(2)
    c.step(dt)
    x, t, norm = c.gnn_snapshot(0)
    rng = np.random.default_rng(4)
    theta = rng.normal(size=c.gnn_np).astype(np.float32) * 0.3
    c.gnn_params(theta)
    g, loss = c.gnn_grad(0)
    nb, A = _graph_from_cloud(c)
    out, _ = oracle.gnn_forward_backward(nb, x.astype(np.float64), theta.astype(np.float64), 64, 1)
    e = out - t
    loss_ref = e @ A @ e / norm
    say(f"one-layer loss gpu {loss:.6f} oracle {loss_ref:.6f}")
    assert abs(loss - loss_ref) / loss_ref < 2e-2


def test_gnn_residual_loss_gradient_matches_oracle(oracle):
    """Residual-norm training loss |A e|^2 / |A d|^2: value and gradient against the oracle
    network driven with dloss/dout = 2 A A e / |A d|^2 (the Laplacian is symmetric)."""
    say("gnn residual-loss gradient vs oracle")
    c, dt = _plasma_cloud(n=(8, 6, 4), ppc=20)
    c.gnn_init(layers=3, width=64, seed=3)
    c.gnn_loss(1)
    c.gnn_record(4)
    for _ in range(3):
        c.step(dt)
    x, t, _ = c.gnn_snapshot(2)
    rng = np.random.default_rng(1)
    theta = c.gnn_params().astype(np.float64) + rng.normal(size=c.gnn_np) * 0.05
    c.gnn_params(theta.astype(np.float32))
    theta = c.gnn_params().astype(np.float64)
    g_gpu, loss_gpu = c.gnn_grad(2)
    nb, A = _graph_from_cloud(c)
    t64 = t.astype(np.float64)
    out, _ = oracle.gnn_forward_backward(nb, x.astype(np.float64), theta, 64, 3)
    e = out - t64
    norm_r = np.sum((A @ t64) ** 2)
    loss_ref = np.sum((A @ e) ** 2) / norm_r
    dout = 2 * (A @ (A @ e)) / norm_r
    _, g_ref = oracle.gnn_forward_backward(nb, x.astype(np.float64), theta, 64, 3, dout)
    cos = g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref))
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    say(f"residual loss gpu {loss_gpu:.6f} oracle {loss_ref:.6f}; gradient cosine {cos:.5f}, rel err {rel:.3e}")
    assert abs(loss_gpu - loss_ref) / loss_ref < 2e-2
    assert cos > 0.995 and rel < 0.1


@pytest.mark.parametrize("kind", [1, 2])
def test_gnn_training_other_losses(kind):
    """Residual-norm (1) and combined (2) losses train down and the guess does not cost
    PCG iterations."""
    say(f"gnn training, loss kind {kind}")
    c, dt = _plasma_cloud()
    c.gnn_init(layers=3, width=64, seed=5)
    c.gnn_loss(kind)
    c.gnn_record(64)
    iters_plain = []
    for k in range(80):
        st = c.step(dt)
        if k >= 16:
            iters_plain.append(st.poisson.iterations)
    l0 = np.mean([c.gnn_eval(i) for i in range(8)])
    first, last = c.gnn_train(600, lr=2e-3, seed=2)
    l1 = np.mean([c.gnn_eval(i) for i in range(8)])
    c.gnn_record(0)
    c.gnn_use(True)
    iters_gnn = [c.step(dt).poisson.iterations for _ in range(40)]
    say(f"loss kind {kind}: before {l0:.4f} after {l1:.4f}; PCG iterations warm start {np.mean(iters_plain):.2f}, "
        f"learned guess {np.mean(iters_gnn):.2f}")
    assert l1 < l0
    assert np.mean(iters_gnn) <= np.mean(iters_plain) + 0.5


def test_extrapolated_warm_start_same_solution_fewer_iterations():
    """poissonGuess extrapolate (2 phi(n-1) - phi(n-2)) changes only where PCG starts: the
    fields agree with the previous-phi warm start to the solver tolerance.  On this
    shot-noise-dominated plasma the step-to-step change of phi is not smooth, so neither
    start saves iterations over the other (within one iteration)."""
    say("extrapolated warm start")
    from picfoam_amd import cases, picmi
    res = {}
    for guess in ("previous", "extrapolate"):
        mesh, species, models, ex = cases.poisson_plasma(ncell=16, ppc=8, guess=guess)
        c = picmi.Cloud(mesh, species, models, ex["bc"])
        c.set_field(picmi.FIELD_SPACE_CHARGE, np.full(ex["n_cells"], ex["space_charge"]))
        c.init_maxwellian(0, 8, ex["T_K"], seed=4)
        its = [c.step(ex["dt"]).poisson.iterations for _ in range(12)]
        res[guess] = (np.array(its), c.get_field(picmi.FIELD_PHI).copy())
        c.close()
    it_p, phi_p = res["previous"]
    it_e, phi_e = res["extrapolate"]
    say(f"iterations previous {it_p.mean():.2f} extrapolate {it_e.mean():.2f}")
    assert np.linalg.norm(phi_e - phi_p) <= 1e-6 * np.linalg.norm(phi_p)
    assert it_e[3:].mean() <= it_p[3:].mean() + 1.0
