"""Pins the far-field oracle (oracle/far_oracle.py), CPU only:
  - the hand-written MLP gradient against central finite differences (float64);
  - the exact pair scorer against the reference-restating C++ oracle: the uncut Yukawa force on
    subunit A's charged beads with subunit B present minus with A alone, per bead;
  - Newton's third law and the torque balance between the two views of a pair;
  - the body frame and pose features: invariant under a global rotation of both subunits;
  - the library's template frame (frame beads, charged-bead body coordinates) against it.
"""
import os
import sys

import numpy as np
import pytest

from oracle_lib import REPO, TEMPLATE, Oracle
from souffle_states import random_rotation

sys.path.insert(0, os.path.join(REPO, "oracle"))
import far_oracle as fo  # noqa: E402


def _template_cols(path):
    tok = open(path).read().split()
    n = int(tok[14])
    arr = np.array(tok[23:23 + 8 * n], dtype=float).reshape(n, 8)
    return arr[:, 1:4].copy(), arr[:, 5].copy(), arr[:, 6].copy()


def _centre(xyz):
    return int(np.argmin(np.linalg.norm(xyz - xyz.mean(axis=0), axis=1)))


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.2, size=fo.N_PARAMS)
    x = rng.normal(size=(7, fo.N_IN))
    y = rng.normal(size=(7, fo.N_OUT))
    net = fo.MLP(p)
    loss, g = net.loss_and_grad(x, y)
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in
                          [(fo.W1, fo.B1), (fo.B1, fo.W2), (fo.W2, fo.B2), (fo.B2, fo.W3), (fo.W3, fo.B3),
                           (fo.B3, fo.W4), (fo.W4, fo.B4), (fo.B4, fo.N_PARAMS)]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (fo.MLP(pp).loss_and_grad(x, y)[0] - fo.MLP(pm).loss_and_grad(x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    assert loss == pytest.approx(np.mean((net.forward(x) - y) ** 2))


@pytest.mark.parametrize("c_mM,d,seed", [(10.0, 22.0, 1), (100.0, 21.0, 2), (10.0, 40.0, 3)])
def test_pair_exact_matches_reference_oracle(c_mM, d, seed):
    xyz, q, sig = _template_cols(TEMPLATE)
    c = _centre(xyz)
    fp1, fp2, R0 = fo.template_frame(xyz, c)
    qi = np.nonzero(q)[0]
    qbody = (xyz[qi] - xyz[c]) @ R0
    rng = np.random.default_rng(seed)
    n = rng.normal(size=3)
    dvec = d * n / np.linalg.norm(n)
    Q = random_rotation(rng)
    # lab coordinates: A as in the template, B = rigid copy at the pose (dvec, Q) in A's frame
    xa = xyz.copy()
    xb = xyz[c] + R0 @ dvec + (xyz - xyz[c]) @ (R0 @ Q @ R0.T).T
    o2 = Oracle(2, template=TEMPLATE, C_uM=1.0, c_mM=c_mM, ecut_c=0.0, L=400.0, B=1)
    o1 = Oracle(1, template=TEMPLATE, C_uM=0.5, c_mM=c_mM, ecut_c=0.0, L=400.0, B=1)
    assert abs(o1.box - o2.box) < 1e-9 * o2.box and o2.box > 2 * (d + 10)
    o2.set_state(np.concatenate([xa, xb]))
    o1.set_state(xa)
    f_alone = o1.forces()[3]
    f_pair = (o2.forces()[3][:len(xyz)] - f_alone)[qi]
    F, T, f = fo.pair_exact(qbody, q[qi], sig[qi], o2.lb, dvec, Q, o2.kappa)
    # the difference carries the round-off of A's own (much larger) intra-subunit forces
    tol = 1e-9 * np.abs(f).max() + 1e-14 * np.abs(f_alone).max()
    assert np.abs(f @ R0.T - f_pair).max() <= tol
    assert np.allclose(R0 @ F, f_pair.sum(axis=0), rtol=0, atol=10 * tol)
    torque = np.cross(xyz[qi] - xyz[c], f_pair).sum(axis=0)
    assert np.allclose(R0 @ T, torque, rtol=0, atol=100 * tol)


def test_pair_exact_action_reaction():
    xyz, q, sig = _template_cols(TEMPLATE)
    c = _centre(xyz)
    _, _, R0 = fo.template_frame(xyz, c)
    qi = np.nonzero(q)[0]
    qbody = (xyz[qi] - xyz[c]) @ R0
    rng = np.random.default_rng(7)
    for _ in range(3):
        dvec = rng.normal(size=3)
        dvec *= 25.0 / np.linalg.norm(dvec)
        Q = random_rotation(rng)
        kappa = 0.4
        F, T, _ = fo.pair_exact(qbody, q[qi], sig[qi], 0.701, dvec, Q, kappa)
        # the same pair seen from B: A at -Q^T d, orientation Q^T
        Fb, Tb, _ = fo.pair_exact(qbody, q[qi], sig[qi], 0.701, -Q.T @ dvec, Q.T, kappa)
        # in A's frame: F_B = -F_A, tau_B = -tau_A - (c_A - c_B) x F_A with c_A - c_B = -d
        np.testing.assert_allclose(Q @ Fb, -F, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(Q @ Tb, -T + np.cross(dvec, F), rtol=1e-9, atol=1e-13)


def test_features_invariant_under_global_rotation():
    xyz, q, _ = _template_cols(TEMPLATE)
    c = _centre(xyz)
    fp1, fp2, R0 = fo.template_frame(xyz, c)
    rng = np.random.default_rng(11)
    Q = random_rotation(rng)
    dvec = np.array([12.0, -17.0, 9.0])
    xa = xyz.copy()
    xb = xyz[c] + R0 @ dvec + (xyz - xyz[c]) @ (R0 @ Q @ R0.T).T

    def feats(a, b):
        Ra = fo.body_frame(a[fp1] - a[c], a[fp2] - a[c])
        Rb = fo.body_frame(b[fp1] - b[c], b[fp2] - b[c])
        return fo.pose_features(Ra.T @ (b[c] - a[c]), Ra.T @ Rb, 20.0, 0.5)

    f0 = feats(xa, xb)
    np.testing.assert_allclose(f0, fo.pose_features(dvec, Q, 20.0, 0.5), atol=1e-12)
    G = random_rotation(rng)
    shift = np.array([3.0, -1.0, 2.0])
    np.testing.assert_allclose(feats(xa @ G.T + shift, xb @ G.T + shift), f0, atol=1e-12)


def test_library_template_frame():
    import souffle_gpu as sg
    t = sg.Template(TEMPLATE)
    xyz, q, _ = _template_cols(TEMPLATE)
    c = _centre(xyz)
    assert c == t.centre
    fp1, fp2, R0 = fo.template_frame(xyz, c)
    lfp1, lfp2, qbody = t.frame()
    assert (lfp1, lfp2) == (fp1, fp2)
    np.testing.assert_allclose(qbody, (xyz[q != 0] - xyz[c]) @ R0, atol=1e-12)
