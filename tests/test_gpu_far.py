"""The far-field model of the uncut Yukawa on the GPU, against the CPU oracle
(oracle/far_oracle.py) and the exact far field.

Tolerances: features and scaled targets of the generated poses 1e-5 relative (stored FP32);
loss and gradients of the FP32 training kernels 1e-4 norm-wise against the float64 oracle;
FP32 forward 1e-5; BF16 MFMA inference 3e-2 norm-wise (bf16 keeps 8 mantissa bits; FP32
accumulation).
"""
import os
import sys

import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import REPO, TEMPLATE, rel_norm

sys.path.insert(0, os.path.join(REPO, "oracle"))
import far_oracle as fo  # noqa: E402

pytestmark = pytest.mark.gpu

RF = 20.0
C_LO, C_HI = 5.0, 50.0


def _template_cols(path):
    tok = open(path).read().split()
    n = int(tok[14])
    arr = np.array(tok[23:23 + 8 * n], dtype=float).reshape(n, 8)
    return arr[:, 1:4].copy(), arr[:, 5].copy(), arr[:, 6].copy()


@pytest.fixture(scope="module")
def template():
    return sg.Template(TEMPLATE)


@pytest.fixture(scope="module")
def trained(template):
    m = sg.FarModel(template, RF, C_LO, C_HI, seed=2)
    rep = m.train(n_samples=1000000, epochs=20, batch=4096, lr=2e-3, seed=77)
    print("training report:", rep)
    return m, rep


def test_samples_match_oracle(template):
    m = sg.FarModel(template, RF, C_LO, C_HI, seed=1)
    x, y, pose = m.samples(257, 9, C_LO, C_HI)
    _, _, qbody = template.frame()
    xyz, q, sig = _template_cols(TEMPLATE)
    qi = np.nonzero(q)[0]
    lb = 0.701
    for r in range(len(x)):
        dvec, Q, kappa, S = pose[r, :3], pose[r, 3:12].reshape(3, 3), pose[r, 12], pose[r, 14]
        np.testing.assert_allclose(Q @ Q.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Q) - 1) < 1e-12
        assert np.linalg.norm(dvec) >= RF * (1 - 1e-12)
        np.testing.assert_allclose(x[r], fo.pose_features(dvec, Q, RF, kappa), rtol=1e-6, atol=1e-7)
        F, T, _ = fo.pair_exact(qbody, q[qi], sig[qi], lb, dvec, Q, kappa)
        assert abs(S - fo.far_scale(lb, qbody, dvec, Q, kappa)) <= 1e-12 * S
        yt = np.concatenate([F / S, T / (S * template.r_charged)])
        assert rel_norm(y[r], yt) < 1e-5, r
    # salts inside the range, both distance regimes present
    c = pose[:, 13]
    assert c.min() >= C_LO * (1 - 1e-12) and c.max() <= C_HI * (1 + 1e-12)
    d = np.linalg.norm(pose[:, :3], axis=1)
    assert (d < 3 * RF).any() and (d > 3 * RF).any()


def test_training_gradients_match_oracle(template):
    m = sg.FarModel(template, RF, C_LO, C_HI, seed=4)
    x, _, _ = m.samples(300, 3, C_LO, C_HI)
    y = np.random.default_rng(5).normal(size=(300, 6)).astype(np.float32)
    loss, g = m.grad(x, y)
    p, _, _ = m.params()
    ref_loss, ref_g = fo.MLP(p.astype(np.float64)).loss_and_grad(x.astype(np.float64), y.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    assert rel_norm(g, ref_g) < 1e-4
    # every parameter block carries gradient
    for a, b in [(fo.W1, fo.B1), (fo.W2, fo.B2), (fo.W3, fo.B3), (fo.W4, fo.B4), (fo.B4, fo.N_PARAMS)]:
        assert rel_norm(g[a:b], ref_g[a:b]) < 1e-4, (a, b)


def test_inference_fp32_and_bf16_mfma(template):
    m = sg.FarModel(template, RF, C_LO, C_HI, seed=6)
    x, _, _ = m.samples(1000, 12, C_LO, C_HI)  # not a multiple of the 128-row tile
    p, _, _ = m.params()
    ref = fo.MLP(p.astype(np.float64)).forward(x.astype(np.float64))
    y32 = m.infer(x, bf16_mfma=False)
    y16 = m.infer(x, bf16_mfma=True)
    assert rel_norm(y32, ref) < 1e-5
    assert rel_norm(y16, ref) < 3e-2
    # row by row (no row or column mix-up in the MFMA tiling)
    err = np.linalg.norm(y16 - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-6)
    assert np.percentile(err, 99) < 0.1


def test_training_fits_held_out_poses(trained):
    m, rep = trained
    assert rep["steps"] > 0 and rep["seconds"] > 0
    assert rep["val_rel_err_force"] < 0.05 and rep["val_rel_err_torque"] < 0.05, rep
    # the trained network through the MFMA path stays close to the FP32 forward
    x, _, _ = m.samples(512, 99, C_LO, C_HI)
    assert rel_norm(m.infer(x, True), m.infer(x, False)) < 3e-2


def _engine(template, c_mM, seed=3, ns=200):
    p = template.params(n_subunits=ns, C_uM=800.0, c_mM=c_mM, elj_att=2.1, dt=0.005, K=298.0, ecut_c=0.0)
    p.far_rf = RF
    pos, vel = template.make_state(ns, p.box, p.T, seed)
    s = sg.System(template, ns, p)
    s.set_state(pos, vel)
    return s, pos, vel, p


def test_model_gains_over_dropping_the_far_field(template, trained):
    """What the network adds: the error of the force with the far field left out against the
    error with the network's far field, both against the exact force."""
    m, _ = trained
    s, pos, vel, p = _engine(template, 10.0)
    f_exact = s.forces()[0]
    s.far_drop(True)
    f_none = s.forces()[0]
    s.far_drop(False)
    s.set_far_model(m, tol=1.0)
    f_model = s.forces()[0]
    err_none = rel_norm(f_none, f_exact)
    err_model = rel_norm(f_model, f_exact)
    assert err_none > 0
    # measured: the network's far field is ~6% off per bead, the dropped one 100%
    assert err_model < err_none / 8, (err_model, err_none)
    # the audit measures the model's error relative to the far field it replaces: the forces
    # differ only in the far field, so that is err_model / err_none at this state
    a = s.far_audit()
    assert abs(a - err_model / err_none) <= 1e-6 * a, (a, err_model / err_none)
    st = s.far_stats()
    print("far-field relative error", a, "total-force relative error", st["last_err_force"], "dropped", err_none)
    assert abs(st["last_err_force"] - err_model) <= 1e-3 * err_model


def test_model_audit_falls_back(template, trained):
    m, _ = trained
    s, pos, vel, p = _engine(template, 10.0)
    s.set_far_model(m, tol=1e-14)
    s.step(100)
    st = s.far_stats()
    assert st["audits"] == 1 and st["fell_back"] == 1 and st["using_model"] == 0
    # the failed chunk was run again with the exact far field: the exact run's trajectory
    ex, _, _, _ = _engine(template, 10.0)
    ex.step(100)
    assert all(np.array_equal(a, b) for a, b in zip(s.get_state(), ex.get_state()))
    s2, _, _, _ = _engine(template, 10.0)
    s2.set_far_model(m, tol=1.0)
    s2.step(200)
    st2 = s2.far_stats()
    assert st2["audits"] == 2 and st2["fell_back"] == 0 and st2["using_model"] == 1
    assert st2["max_err"] < 1.0 and st2["max_err_force"] < 1e-3


def test_corrupted_model_trips_the_default_tolerance(template, trained):
    """A model with perturbed weights fails the far-field audit at tol = 1e-3, although the same
    difference is far below 1e-3 of the total force at c = 10 mM (where an audit against the
    total force could never fail)."""
    m, _ = trained
    p0, mu, sd = m.params()
    bad = sg.FarModel(template, RF, C_LO, C_HI, seed=2)
    bad.set_params(p0 * np.float32(1.3), mu, sd)
    s, _, _, _ = _engine(template, 10.0)
    s.set_far_model(bad, tol=1e-3)
    s.step(100)
    st = s.far_stats()
    assert st["fell_back"] == 1 and st["max_err"] > 1e-3, st
    assert st["max_err_force"] < 1e-3, st


def test_model_runs_deterministic(template, trained):
    m, _ = trained
    out = []
    for _ in range(2):
        s, _, _, _ = _engine(template, 20.0)
        s.set_far_model(m, tol=1.0)
        s.step(150)
        out.append(s.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_model_rejects_salt_outside_training_range(template, trained):
    m, _ = trained
    s, _, _, _ = _engine(template, 100.0, ns=60)
    with pytest.raises(sg.SfError):
        s.set_far_model(m, tol=1e-3)
