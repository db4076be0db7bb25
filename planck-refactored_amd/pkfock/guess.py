"""Learned initial density for water clusters in the 6-31G* basis.

The density matrix of a cluster is assembled from blocks, each predicted in the
molecular frame of the waters involved and rotated to the laboratory frame:

* intra-water blocks P_AA (19 x 19, symmetric: 190 numbers) from the
  environment of water A seen in A's frame;
* inter-water blocks P_AB (19 x 19) for O-O distances below 6 Angstrom from the
  position of B's atoms in A's frame, of A's atoms in B's frame and the two
  environments; the prediction for (A, B) and the transpose of the one for
  (B, A) are averaged, so the assembled matrix is symmetric.

Frame of a water: origin O, e1 along the H-O-H bisector, e2 along H1 - H2
(orthogonalised), e3 = e1 x e2.  Function-space rotations: s -> 1, p -> R^T,
Cartesian d -> the 6 x 6 map of the normalised quadratic monomials (component
norms 1/sqrt(3) for xx, yy, zz).  Function order per water (basis order of the
reference, src/basis/basis.cpp:20-37): O s s p s p d (15), H s s, H s s.

The two networks (pkfock.mlp.Net, hand-written HIP kernels) are trained on
converged RHF densities produced by this engine's own SCF.  At SCF time the
learned density is used only when its first DIIS error is not larger than the
core guess's (pk_scf_options.guess_fallback); otherwise the core guess is used
(the reference's plain guess, src/scf/scf.cpp:33-45).  No counterpart in the
reference.
"""
import numpy as np

from .molecules import ANGSTROM_TO_BOHR

NF_O, NF_H = 15, 2
NFW = NF_O + 2 * NF_H  # 19 functions per water
CUTOFF = 6.0  # Angstrom
RADIAL = np.linspace(1.0, 6.0, 8)
SIGMA = 0.5
N_ENV = 2 * 8 * 10  # species x radial x (1, u, u u^T upper)
N_ENV_PAIR = 2 * 8 * 4  # species x radial x (1, u)
N_INTRA_IN = N_ENV
N_INTER_IN = 9 + 9 + 3 + 2 * N_ENV_PAIR
N_INTRA_OUT = NFW * (NFW + 1) // 2
N_INTER_OUT = NFW * NFW
_TRIU = np.triu_indices(NFW)

# layout of one water's 19 functions: (kind, offset) with kind s / p / d
_LAYOUT = [("s", 0), ("s", 1), ("p", 2), ("s", 5), ("p", 6), ("d", 9), ("s", 15), ("s", 16), ("s", 17), ("s", 18)]
_DPAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
_DNORM = np.array([1 / np.sqrt(3.0), 1.0, 1.0, 1 / np.sqrt(3.0), 1.0, 1 / np.sqrt(3.0)])


def water_frames(X):
    """X: (3n, 3) Angstrom, atoms O H H per water.  Returns origins (n, 3) and
    R (n, 3, 3) with the frame axes as columns (r_lab = O + R r_local)."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3, 3)
    O, H1, H2 = X[:, 0], X[:, 1], X[:, 2]
    e1 = (H1 - O) + (H2 - O)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = H1 - H2
    e2 -= (e2 * e1).sum(1, keepdims=True) * e1
    e2 /= np.linalg.norm(e2, axis=1, keepdims=True)
    e3 = np.cross(e1, e2)
    return O, np.stack([e1, e2, e3], axis=2)


def d_transform(R):
    """6 x 6 map M with phi_local = M phi_lab for normalised Cartesian d
    functions, local coordinates x' = R^T x."""
    M = np.zeros((6, 6))
    for a, (i, j) in enumerate(_DPAIRS):
        for b, (k, l) in enumerate(_DPAIRS):
            v = R[k, i] * R[l, j]
            if k != l:
                v += R[l, i] * R[k, j]
            M[a, b] = _DNORM[a] * v / _DNORM[b]
    return M


def water_transform(R):
    """19 x 19 map M of one water's functions: phi_local = M phi_lab."""
    M = np.zeros((NFW, NFW))
    for kind, o in _LAYOUT:
        if kind == "s":
            M[o, o] = 1.0
        elif kind == "p":
            M[o:o + 3, o:o + 3] = R.T
        else:
            M[o:o + 6, o:o + 6] = d_transform(R)
    return M


def _env(X, O, R, a, full):
    """Environment of water a in its frame: smooth-cutoff Gaussian radial bins
    times angular factors of the unit vector, per species (O, H)."""
    nw = len(O)
    out = np.zeros((2, 8, 10 if full else 4))
    Xw = np.asarray(X).reshape(nw, 3, 3)
    for b in range(nw):
        if b == a:
            continue
        for k in range(3):
            d = Xw[b, k] - O[a]
            r = np.linalg.norm(d)
            if r >= CUTOFF or r < 1e-8:
                continue
            u = R[a].T @ d / r
            fc = 0.5 * (np.cos(np.pi * r / CUTOFF) + 1.0)
            g = np.exp(-((r - RADIAL) ** 2) / (2 * SIGMA * SIGMA)) * fc
            ang = [1.0, u[0], u[1], u[2]]
            if full:
                ang += [u[0] * u[0], u[1] * u[1], u[2] * u[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]]
            out[0 if k == 0 else 1] += np.outer(g, ang)
    return out.ravel()


def pairs_within(O):
    n = len(O)
    out = []
    for a in range(n):
        for b in range(n):
            if a != b and np.linalg.norm(O[a] - O[b]) < CUTOFF:
                out.append((a, b))
    return out


def features(X):
    """Intra features (n, N_INTRA_IN), ordered pairs (a, b) within the cutoff and
    their inter features (npairs, N_INTER_IN), frames and transforms."""
    O, R = water_frames(X)
    nw = len(O)
    Xw = np.asarray(X, dtype=np.float64).reshape(nw, 3, 3)
    intra = np.stack([_env(X, O, R, a, True) for a in range(nw)]) if nw else np.zeros((0, N_INTRA_IN))
    envp = [_env(X, O, R, a, False) for a in range(nw)]
    pairs = pairs_within(O)
    inter = np.zeros((len(pairs), N_INTER_IN))
    for i, (a, b) in enumerate(pairs):
        pb = (R[a].T @ (Xw[b] - O[a]).T).T.ravel()  # B's atoms in A's frame
        pa = (R[b].T @ (Xw[a] - O[b]).T).T.ravel()  # A's atoms in B's frame
        r = np.linalg.norm(O[b] - O[a])
        inter[i] = np.concatenate([pb, pa, [r, 1.0 / r, np.exp(-r)], envp[a], envp[b]])
    return intra, pairs, inter, O, R


def local_blocks(P, R, pairs):
    """Lab density (19 n x 19 n) -> local intra blocks (n, 190) and inter
    blocks (npairs, 361): P_local = M_a^-T P_ab M_b^-1."""
    nw = len(R)
    Minv = [np.linalg.inv(water_transform(R[a])) for a in range(nw)]
    intra = np.zeros((nw, N_INTRA_OUT))
    for a in range(nw):
        blk = P[a * NFW:(a + 1) * NFW, a * NFW:(a + 1) * NFW]
        intra[a] = (Minv[a].T @ blk @ Minv[a])[_TRIU]
    inter = np.zeros((len(pairs), N_INTER_OUT))
    for i, (a, b) in enumerate(pairs):
        blk = P[a * NFW:(a + 1) * NFW, b * NFW:(b + 1) * NFW]
        inter[i] = (Minv[a].T @ blk @ Minv[b]).ravel()
    return intra, inter


def assemble(intra, pairs, inter, R):
    """Local blocks -> symmetric lab density: P_ab = M_a^T P_local M_b; the two
    orientations of an inter block are averaged."""
    nw = len(R)
    M = [water_transform(R[a]) for a in range(nw)]
    P = np.zeros((nw * NFW, nw * NFW))
    for a in range(nw):
        L = np.zeros((NFW, NFW))
        L[_TRIU] = intra[a]
        L = L + np.triu(L, 1).T
        P[a * NFW:(a + 1) * NFW, a * NFW:(a + 1) * NFW] = M[a].T @ L @ M[a]
    acc = {}
    for i, (a, b) in enumerate(pairs):
        blk = M[a].T @ inter[i].reshape(NFW, NFW) @ M[b]
        key, val = ((a, b), blk) if a < b else ((b, a), blk.T)
        acc.setdefault(key, []).append(val)
    for (a, b), vals in acc.items():
        blk = sum(vals) / len(vals)
        P[a * NFW:(a + 1) * NFW, b * NFW:(b + 1) * NFW] = blk
        P[b * NFW:(b + 1) * NFW, a * NFW:(a + 1) * NFW] = blk.T
    return P


class LearnedGuess:
    """Two device networks (intra, inter) with output standardisation."""

    def __init__(self, hidden=128, seed=0, max_batch=65536):
        import torch
        # torch's HIP runtime is initialised before the engine's libraries run
        # more work in this process (device buffers come from torch)
        torch.cuda.init()
        from .mlp import Net
        self.intra = Net(N_INTRA_IN, N_INTRA_OUT, hidden, max_batch, seed)
        self.inter = Net(N_INTER_IN, N_INTER_OUT, hidden, max_batch, seed + 1)
        # the networks run on torch's current stream, so the batches that torch
        # gathers (Xs[idx]) are complete before a kernel reads them
        self.bind_stream()
        self.stats = None

    def bind_stream(self, stream=None):
        """Run both networks on `stream` (default: torch's current stream)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        self.intra.set_stream(s.cuda_stream)
        self.inter.set_stream(s.cuda_stream)

    def fit(self, samples, epochs=200, batch=512, lr=2e-3, log=None, seed=0):
        """samples: list of (X Angstrom, converged P).  Standardises inputs and
        outputs, then Adam on mini-batches, all on the device."""
        import torch
        dev = torch.device("cuda:0")
        self.bind_stream()
        XI, YI, XE, YE = [], [], [], []
        for X, P in samples:
            fi, pairs, fe, O, R = features(X)
            li, le = local_blocks(P, R, pairs)
            XI.append(fi); YI.append(li); XE.append(fe); YE.append(le)
        XI, YI, XE, YE = (np.concatenate(v) for v in (XI, YI, XE, YE))
        self.stats = {k: (v.mean(0), v.std(0) + 1e-6) for k, v in (("xi", XI), ("yi", YI), ("xe", XE), ("ye", YE))}
        rng = np.random.default_rng(seed)
        hist = []
        for net, X, Y, kx, ky in ((self.intra, XI, YI, "xi", "yi"), (self.inter, XE, YE, "xe", "ye")):
            Xs = torch.tensor(((X - self.stats[kx][0]) / self.stats[kx][1]).astype(np.float32), device=dev)
            Ys = torch.tensor(((Y - self.stats[ky][0]) / self.stats[ky][1]).astype(np.float32), device=dev)
            n = Xs.shape[0]
            for ep in range(epochs):
                perm = torch.tensor(rng.permutation(n), device=dev)
                tot = 0.0
                for s in range(0, n, batch):
                    idx = perm[s:s + batch]
                    xb, yb = Xs[idx].contiguous(), Ys[idx].contiguous()
                    tot += net.loss_grad(xb, yb) * len(idx)
                    net.adam_step(lr * (0.1 ** (ep / epochs)))
                hist.append((kx, ep, tot / n))
                if log and (ep % 20 == 0 or ep == epochs - 1):
                    log("%s epoch %d loss %.4e" % ("intra" if kx == "xi" else "inter", ep, tot / n))
        torch.cuda.synchronize()
        return hist

    def predict_blocks(self, X):
        import torch
        dev = torch.device("cuda:0")
        self.bind_stream()
        fi, pairs, fe, O, R = features(X)
        out = []
        for net, F, kx, ky in ((self.intra, fi, "xi", "yi"), (self.inter, fe, "xe", "ye")):
            if len(F) == 0:
                out.append(np.zeros((0, net.n_out)))
                continue
            xs = torch.tensor(((F - self.stats[kx][0]) / self.stats[kx][1]).astype(np.float32), device=dev)
            ys = torch.empty((len(F), net.n_out), dtype=torch.float32, device=dev)
            net.forward(xs, ys)
            torch.cuda.synchronize()
            out.append(ys.cpu().numpy().astype(np.float64) * self.stats[ky][1] + self.stats[ky][0])
        return out[0], pairs, out[1], R

    def predict(self, X):
        intra, pairs, inter, R = self.predict_blocks(X)
        return assemble(intra, pairs, inter, R)

    def state(self):
        return {"intra": self.intra.params(), "inter": self.inter.params(),
                **{"stat_%s_%d" % (k, i): v[i] for k, v in self.stats.items() for i in range(2)}}

    @classmethod
    def pretrained(cls):
        """The network trained by pkfock.train_guess (data/water_631gs_guess.npz)."""
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "water_631gs_guess.npz")
        st = dict(np.load(path))
        m = cls(hidden=128)
        m.load_state(st)
        return m

    def load_state(self, st):
        self.intra.set_params(st["intra"])
        self.inter.set_params(st["inter"])
        self.stats = {k: (st["stat_%s_0" % k], st["stat_%s_1" % k]) for k in ("xi", "yi", "xe", "ye")}


def scf_with_guess(mol, model, **scf):
    """RHF from the learned density with the plain-guess fallback: the SCF
    driver builds the Fock matrix of the learned density and of the core guess,
    and starts from the one with the smaller first DIIS error (one extra Fock
    build).  Returns (result, "learned" | "core")."""
    X = mol.xyz / ANGSTROM_TO_BOHR
    Pg = model.predict(X)
    r = mol.rhf(guess=Pg, guess_fallback=True, **scf)
    return r, ("learned" if r["guess_used"] else "core")
