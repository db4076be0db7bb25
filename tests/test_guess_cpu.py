"""Host-side pieces of the learned guess (no GPU): frames, function-space
rotations, features and block round trips; and the network oracle pinned by
finite differences."""
import os
import sys

import numpy as np
import pytest

from pkfock import guess as G
from pkfock.molecules import water_cluster

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def _rot(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_d_transform_rotates_functions():
    # phi_local(r) = c_ij x'_i x'_j with x' = R^T x must equal sum_kl M[ij, kl] phi_kl(r)
    R = _rot(3)
    M = G.d_transform(R)
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.normal(size=3)
        lab = np.array([G._DNORM[a] * x[i] * x[j] for a, (i, j) in enumerate(G._DPAIRS)])
        xl = R.T @ x
        loc = np.array([G._DNORM[a] * xl[i] * xl[j] for a, (i, j) in enumerate(G._DPAIRS)])
        assert np.abs(M @ lab - loc).max() < 1e-13
    assert np.abs(G.d_transform(R) @ G.d_transform(R.T) - np.eye(6)).max() < 1e-13
    assert np.abs(G.d_transform(np.eye(3)) - np.eye(6)).max() < 1e-15


def test_frames_are_rotation_covariant():
    Z, X = water_cluster(5, n=4)
    O, R = G.water_frames(X)
    for a in range(4):
        assert np.abs(R[a].T @ R[a] - np.eye(3)).max() < 1e-12 and np.linalg.det(R[a]) > 0
    Q = _rot(7)
    fi, pairs, fe, _, _ = G.features(X)
    fi2, pairs2, fe2, _, _ = G.features(X @ Q.T + np.array([1.0, -2.0, 0.5]))
    assert pairs == pairs2
    assert np.abs(fi - fi2).max() < 1e-10 and np.abs(fe - fe2).max() < 1e-10


def test_block_round_trip_and_covariance():
    Z, X = water_cluster(11, n=5)
    fi, pairs, fe, O, R = G.features(X)
    n = 5 * G.NFW
    rng = np.random.default_rng(2)
    A = rng.normal(size=(n, n))
    P = A + A.T
    # only blocks within the cutoff survive the round trip
    mask = np.zeros((n, n), bool)
    for a in range(5):
        mask[a * 19:(a + 1) * 19, a * 19:(a + 1) * 19] = True
    for a, b in pairs:
        mask[a * 19:(a + 1) * 19, b * 19:(b + 1) * 19] = True
    P = np.where(mask, P, 0.0)
    li, le = G.local_blocks(P, R, pairs)
    assert np.abs(G.assemble(li, pairs, le, R) - P).max() < 1e-11
    # local blocks of a rigidly rotated density are unchanged: rotate the lab
    # functions with the transforms of a global rotation Q
    Q = _rot(9)
    X2 = X @ Q.T
    _, pairs2, _, _, R2 = G.features(X2)
    Mq = np.zeros((n, n))
    for a in range(5):
        Mq[a * 19:(a + 1) * 19, a * 19:(a + 1) * 19] = G.water_transform(Q.T)  # phi_new = M phi_old
    P2 = np.linalg.inv(Mq).T @ P @ np.linalg.inv(Mq)
    li2, le2 = G.local_blocks(P2, R2, pairs2)
    assert np.abs(li2 - li).max() < 1e-10 and np.abs(le2 - le).max() < 1e-10


@pytest.mark.parametrize("seed", [0, 1])
def test_mlp_oracle_gradient_by_finite_differences(seed):
    rng = np.random.default_rng(seed)
    n_in, hid, n_out, b = 5, 7, 3, 6
    n_par = n_in * hid + hid + hid * hid + hid + hid * n_out + n_out
    p = rng.normal(size=n_par) * 0.4
    X, T = rng.normal(size=(b, n_in)), rng.normal(size=(b, n_out))
    loss, g = mlp_ref.loss_grad(p, X, T, n_in, hid, n_out)
    for i in rng.choice(n_par, 25, replace=False):
        e = np.zeros(n_par)
        e[i] = 1e-6
        fd = (mlp_ref.loss_grad(p + e, X, T, n_in, hid, n_out)[0] - mlp_ref.loss_grad(p - e, X, T, n_in, hid, n_out)[0]) / 2e-6
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    y, _ = mlp_ref.forward(p, X, n_in, hid, n_out)
    assert abs(loss - ((y - T) ** 2).mean()) < 1e-14


def test_mlp_oracle_adam_first_step_is_signed_lr():
    # with zero moments the first bias-corrected Adam step is lr * g / (|g| + eps)
    g = np.array([0.3, -2.0, 1e-3])
    p, m1, m2 = mlp_ref.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.abs(p + 0.01 * g / (np.abs(g) + 1e-8)).max() < 1e-12


def test_mlp_binding_refuses_host_and_misshapen_tensors():
    """The network kernels take raw device pointers: the binding must refuse a
    host tensor, another dtype, a wrong width, a strided view and row-count
    mismatches before any pointer reaches the library."""
    import types

    import torch
    from pkfock.mlp import Net, PkMlpError

    fake = types.SimpleNamespace(max_batch=8)
    check = Net._device_rows
    with pytest.raises(PkMlpError, match="device"):
        check(fake, torch.zeros(4, 3, dtype=torch.float32), 3, "X")

    class FakeDev:  # stands in for a device tensor's metadata on a CPU-only host
        def __init__(self, t):
            self.t = t
            self.is_cuda = True
            self.dtype, self.shape = t.dtype, t.shape

        def dim(self):
            return self.t.dim()

        def is_contiguous(self):
            return self.t.is_contiguous()

    ok = FakeDev(torch.zeros(4, 3, dtype=torch.float32))
    check(fake, ok, 3, "X")
    check(fake, ok, 3, "X", rows=4)
    with pytest.raises(PkMlpError, match="float32"):
        check(fake, FakeDev(torch.zeros(4, 3, dtype=torch.float64)), 3, "X")
    with pytest.raises(PkMlpError, match="shape"):
        check(fake, ok, 5, "X")
    with pytest.raises(PkMlpError, match="contiguous"):
        check(fake, FakeDev(torch.zeros(3, 4, dtype=torch.float32).t()), 3, "X")
    with pytest.raises(PkMlpError, match="rows"):
        check(fake, ok, 3, "T", rows=5)
    with pytest.raises(PkMlpError, match="max_batch"):
        check(fake, FakeDev(torch.zeros(9, 3, dtype=torch.float32)), 3, "X")
