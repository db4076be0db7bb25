"""The learned-guess network on the device against the CPU oracle
(oracle/mlp_ref.py): forward values, loss, every gradient, Adam steps; and a
small learned guess end to end (trained network, SCF from its density)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def _net(n_in, n_out, hid, seed):
    from pkfock.mlp import Net
    return Net(n_in, n_out, hid, max_batch=4096, seed=seed)


@pytest.mark.parametrize("shape", [(37, 64, 19, 200), (160, 128, 190, 1000), (149, 128, 361, 333)])
def test_forward_and_gradients_match_oracle(shape):
    import torch
    n_in, hid, n_out, b = shape
    rng = np.random.default_rng(n_in)
    net = _net(n_in, n_out, hid, 3)
    p = net.params().astype(np.float64)
    p += rng.normal(size=p.size) * 0.01  # non-zero biases
    net.set_params(p.astype(np.float32))
    p = net.params().astype(np.float64)
    X = rng.normal(size=(b, n_in)).astype(np.float32)
    T = rng.normal(size=(b, n_out)).astype(np.float32)
    dev = torch.device("cuda:0")
    Xd, Td = torch.tensor(X, device=dev), torch.tensor(T, device=dev)
    Y = torch.empty((b, n_out), dtype=torch.float32, device=dev)
    net.forward(Xd, Y)
    torch.cuda.synchronize()
    yref, _ = mlp_ref.forward(p, X, n_in, hid, n_out)
    dy = np.abs(Y.cpu().numpy() - yref).max() / np.abs(yref).max()
    loss = net.loss_grad(Xd, Td)
    lref, gref = mlp_ref.loss_grad(p, X, T, n_in, hid, n_out)
    g = net.grads().astype(np.float64)
    dg = np.abs(g - gref).max() / np.abs(gref).max()
    print("shape %s: forward rel %.2e  loss %.6e vs %.6e  grad rel %.2e" % (shape, dy, loss, lref, dg))
    assert dy < 2e-5 and abs(loss - lref) < 2e-5 * lref and dg < 1e-4


def test_adam_steps_match_oracle():
    import torch
    n_in, hid, n_out, b = 12, 32, 7, 64
    rng = np.random.default_rng(1)
    net = _net(n_in, n_out, hid, 5)
    p = net.params().astype(np.float64)
    X = rng.normal(size=(b, n_in)).astype(np.float32)
    T = rng.normal(size=(b, n_out)).astype(np.float32)
    dev = torch.device("cuda:0")
    Xd, Td = torch.tensor(X, device=dev), torch.tensor(T, device=dev)
    m1, m2 = np.zeros_like(p), np.zeros_like(p)
    for step in range(1, 4):
        net.loss_grad(Xd, Td)
        net.adam_step(1e-2)
        _, g = mlp_ref.loss_grad(p, X, T, n_in, hid, n_out)
        p, m1, m2 = mlp_ref.adam(p, g, m1, m2, step, 1e-2)
    d = np.abs(net.params() - p).max()
    print("params after 3 Adam steps: max diff %.2e" % d)
    assert d < 1e-4


def test_training_reduces_loss():
    import torch
    rng = np.random.default_rng(4)
    X = rng.normal(size=(2048, 8)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) * X[:, 1], X[:, 2] ** 2 - X[:, 3]], 1).astype(np.float32)
    net = _net(8, 2, 64, 0)
    dev = torch.device("cuda:0")
    Xd, Td = torch.tensor(X, device=dev), torch.tensor(T, device=dev)
    first = net.loss_grad(Xd, Td)
    for _ in range(400):
        net.loss_grad(Xd, Td)
        net.adam_step(3e-3)
    last = net.loss_grad(Xd, Td)
    print("loss %.3e -> %.3e" % (first, last))
    assert last < 0.1 * first


def test_learned_guess_end_to_end_small():
    """Train on a few small clusters, then the SCF from the learned density
    reaches the core-guess energy (1e-8 Eh) on a held-out cluster."""
    from pkfock.guess import LearnedGuess, scf_with_guess
    from pkfock.host import Molecule
    from pkfock.molecules import water_cluster
    samples = []
    for s in range(12):
        Z, X = water_cluster(1000 + s, n_min=2, n_max=4)
        r = Molecule(Z, X, "6-31g*").rhf(tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
        samples.append((X, r["P"]))
    model = LearnedGuess(hidden=64, seed=0)
    hist = model.fit(samples, epochs=60, batch=256, lr=3e-3)
    Z, X = water_cluster(5000, n=5)
    mol = Molecule(Z, X, "6-31g*")
    ref = mol.rhf(tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
    res, used = scf_with_guess(mol, model, tol_eri=1e-12, tol_energy=1e-9, tol_density=1e-7)
    print("final losses", hist[59][2], hist[-1][2], "guess used:", used, "iterations", ref["iterations"], "->",
          res["iterations"], "dE %.2e" % (res["energy"] - ref["energy"]))
    assert res["rc"] == 0 and abs(res["energy"] - ref["energy"]) < 1e-8


def test_pretrained_guess_gain():
    """The shipped network (trained on 2-16-water clusters) against the core
    guess on a held-out 16-water cluster: fewer SCF iterations, same energy."""
    from pkfock.guess import LearnedGuess, scf_with_guess
    from pkfock.host import Molecule
    from pkfock.molecules import water_cluster
    model = LearnedGuess.pretrained()
    Z, X = water_cluster(910000, n=16)
    mol = Molecule(Z, X, "6-31g*")
    scf = dict(tol_eri=1e-12, tol_energy=1e-8, tol_density=1e-6)
    core = mol.rhf(**scf)
    lrn, used = scf_with_guess(mol, model, **scf)
    print("16 waters: core %d iterations, learned (%s) %d, first DIIS errors %.3e vs %.3e, dE %.1e" % (
        core["iterations"], used, lrn["iterations"], lrn["guess_diis_error"], lrn["core_diis_error"],
        lrn["energy"] - core["energy"]))
    assert used == "learned" and lrn["iterations"] <= 0.7 * core["iterations"]
    assert abs(lrn["energy"] - core["energy"]) < 1e-8


@pytest.mark.parametrize("seed", [900000, 900001])
def test_pretrained_guess_gain_48_waters(seed):
    """The learned-guess workload at its size: held-out 48-water clusters (912
    functions, seeds outside the training range).  The SCF from the learned
    density (with the first-DIIS-error fallback) needs at least 30 % fewer
    iterations than from the core guess and converges to the same energy
    within 1e-8 Eh."""
    from pkfock.guess import LearnedGuess, scf_with_guess
    from pkfock.host import Molecule
    from pkfock.molecules import water_cluster
    model = LearnedGuess.pretrained()
    Z, X = water_cluster(seed, n=48)
    mol = Molecule(Z, X, "6-31g*")
    scf = dict(tol_eri=1e-12, tol_energy=1e-8, tol_density=1e-6)
    core = mol.rhf(**scf)
    lrn, used = scf_with_guess(mol, model, **scf)
    print("48 waters (seed %d, nb %d): core %d iterations, learned (%s) %d, dE %.1e" % (
        seed, mol.nbasis, core["iterations"], used, lrn["iterations"], lrn["energy"] - core["energy"]))
    assert core["rc"] == 0 and lrn["rc"] == 0
    assert used == "learned" and lrn["iterations"] <= 0.7 * core["iterations"]
    assert abs(lrn["energy"] - core["energy"]) < 1e-8
