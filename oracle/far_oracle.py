"""TEST INFRASTRUCTURE ONLY: CPU restatement (numpy, float64) of the far-field model pieces of
the uncut Yukawa. Only the tests import it, as the checker of the GPU library.

- body_frame / template_frame: the subunit body frame (centre bead, the bead farthest from it,
  the bead farthest from that axis);
- pose_features: the 13 network inputs of a pose (B's centre in A's body frame d, relative
  orientation Q = R_A^T R_B, kappa);
- pair_exact: Yukawa force on A's charged beads from B's, and its torque about A's centre
  bead, for rigid subunits without periodic images (pair term of src/forces.cpp:104-107);
- MLP: the 13-128-128-128-6 SiLU network with the library's packed parameter layout, its
  mean-squared-error loss and gradient (hand-written backward, pinned by finite differences
  in the tests).
"""
import numpy as np

N_IN, N_HID, N_OUT = 13, 128, 6
W1, B1 = 0, N_HID * N_IN
W2 = B1 + N_HID
B2 = W2 + N_HID * N_HID
W3 = B2 + N_HID
B3 = W3 + N_HID * N_HID
W4 = B3 + N_HID
B4 = W4 + N_OUT * N_HID
N_PARAMS = B4 + N_OUT


def body_frame(u1, u2):
    """R with columns e1 = u1/|u1|, e2 = normalised part of u2 normal to e1, e3 = e1 x e2."""
    e1 = u1 / np.linalg.norm(u1)
    w = u2 - np.dot(u2, e1) * e1
    e2 = w / np.linalg.norm(w)
    return np.column_stack([e1, e2, np.cross(e1, e2)])


def template_frame(xyz, centre):
    """Frame beads (first index on ties within 1e-12) and the template's own frame."""
    u = xyz - xyz[centre]
    n2 = np.einsum("ij,ij->i", u, u)
    fp1 = int(np.argmax(n2 > n2.max() - 1e-12))
    e1 = u[fp1] / np.linalg.norm(u[fp1])
    w2 = n2 - (u @ e1) ** 2
    fp2 = int(np.argmax(w2 > w2.max() - 1e-12))
    return fp1, fp2, body_frame(u[fp1], u[fp2])


def pose_features(dvec, Q, rf, kappa):
    d = np.linalg.norm(dvec)
    w = rf / d
    return np.concatenate([dvec / d, [w, w * w, w ** 3], Q[:, 0], Q[:, 1], [kappa]])


def far_scale(lb, qbody, dvec, Q, kappa):
    """Pair-force magnitude lb (kappa + 1/d)/d exp(-kappa d) in the leading (parallel-projection)
    form of the screened sum: times sum_a exp(kappa p_a) sum_b exp(kappa p'_b)."""
    d = np.linalg.norm(dvec)
    n = dvec / d
    qbody = np.asarray(qbody, dtype=float)
    pa = qbody @ n
    pb = -(qbody @ Q.T) @ n
    return lb * (kappa + 1 / d) / d * np.exp(-kappa * d) * np.exp(kappa * pa).sum() * np.exp(kappa * pb).sum()


def pair_exact(qbody, q, sig, lb, dvec, Q, kappa):
    """(F [3], T [3], f [nq, 3]): forces on A's charged beads at qbody from B's at dvec + Q qbody."""
    xa = np.asarray(qbody, dtype=float)
    xb = dvec + xa @ Q.T
    f = np.zeros_like(xa)
    for a in range(len(xa)):
        for b in range(len(xb)):
            r = xa[a] - xb[b]
            rr = np.linalg.norm(r)
            shc = (sig[a] + sig[b]) / 2
            Y = np.exp(kappa * shc) / ((1 + kappa * sig[a] / 2) * (1 + kappa * sig[b] / 2))
            f[a] += r * (q[a] * q[b] * lb * np.exp(-kappa * rr) / rr ** 2) * (kappa + 1 / rr) * Y
    return f.sum(axis=0), np.cross(xa, f).sum(axis=0), f


def _silu(z):
    return z / (1 + np.exp(-z))


def _silu_prime(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


class MLP:
    """Float64 network with the library's packed parameters."""

    def __init__(self, params):
        self.p = np.asarray(params, dtype=np.float64).copy()

    def layers(self):
        p = self.p
        return [(p[W1:B1].reshape(N_HID, N_IN), p[B1:W2]), (p[W2:B2].reshape(N_HID, N_HID), p[B2:W3]),
                (p[W3:B3].reshape(N_HID, N_HID), p[B3:W4]), (p[W4:B4].reshape(N_OUT, N_HID), p[B4:])]

    def forward(self, x, keep=False):
        a = np.asarray(x, dtype=np.float64)
        zs, acts = [], [a]
        for l, (W, b) in enumerate(self.layers()):
            z = a @ W.T + b
            zs.append(z)
            a = _silu(z) if l < 3 else z
            acts.append(a)
        return (a, zs, acts) if keep else a

    def loss_and_grad(self, x, y):
        """Mean over rows and outputs of (yhat - y)^2, and its gradient (packed)."""
        yhat, zs, acts = self.forward(x, keep=True)
        n = yhat.size
        dz = 2 * (yhat - y) / n
        g = np.zeros(N_PARAMS)
        ofs = [(W1, B1, W2), (W2, B2, W3), (W3, B3, W4), (W4, B4, N_PARAMS)]
        layers = self.layers()
        for l in range(3, -1, -1):
            wo, bo, end = ofs[l]
            g[wo:bo] = (dz.T @ acts[l]).ravel()
            g[bo:end] = dz.sum(axis=0)
            if l == 0:
                break
            dz = (dz @ layers[l][0]) * _silu_prime(zs[l - 1])
        return float(np.mean((yhat - y) ** 2)), g
