"""The GPU path against the oracle at the workloads' own sizes, and on a template that exercises the
reference's asymmetric pair rules (src/forces.cpp:117-138, Appendix-A semantics of the survey).

Tolerances as in test_gpu_parity.py: forces and energies 1e-10 norm-wise relative (FP64 path;
the oracle's FP64 mode and the GPU differ by summation order only); trajectories 1e-8 after 100
(2000 subunits) or 200 steps; FP32 pair arithmetic 1e-6 against the oracle's FP64 forces
(measured ~1e-8).

- 2000 trimer_MVM subunits (210,000 beads; C = 800 uM, c = 100 mM): per-term forces on a
  perturbed seeded state, 100-step Nose-Hoover MD and BD trajectories, FP32 pair forces. The
  oracle runs the reference list algorithm at -L 0.3 -B 5 (valid at 100 mM).
- 256 replicas, subunits U[20, 400], salt 50-800 mM, epsilon 1.5-2.5, 280-320 K (FP64 pairs):
  forces of sampled replicas against the oracle with each replica's parameters, and those
  replicas re-run alone bit for bit.
- 5000 subunits, uncut Yukawa (c = 10 mM): the far field of sampled subunits (every subunit pair
  whose centre beads are beyond 20 sigma) against a direct numpy sum of the reference's pair
  term over their charged beads.
- A trimer_MVM variant with type-2 beads of diameter 1.2 and mass 1.3, three type-7 beads per
  subunit (diameter 0.9, mass 0.8), and ordered rows (1, 2) and (2, 7): sigma_i in the attractive
  force, epsilon 50 for type_j == 7, rows applied in one direction only; total force nonzero, so
  the centre-of-mass removal acts.
"""
import numpy as np
import pytest

import replica_sweep
import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz
from template_writer import read_template, write_template

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
NAMES = ["total", "stretch", "bend", "yukawa", "lj"]


def _perturbed_state(t, ns, box, seed, scale=0.05):
    pos, vel = t.make_state(ns, box, 1.0, seed)
    pos = pos + np.random.default_rng(seed).normal(scale=scale, size=pos.shape)
    return pos, vel


def _check_forces(got, ref, tol=1e-10):
    for name, g, r in zip(NAMES, got, ref):
        if np.linalg.norm(r) == 0:
            assert np.linalg.norm(g) == 0, name
            continue
        assert rel_norm(g, r) < tol, (name, rel_norm(g, r))


@pytest.fixture(scope="module")
def tmpl():
    return sg.Template(TEMPLATE)


# ---- 2000 subunits ------------------------------------------------------------------------

@pytest.fixture(scope="module")
def state2000(tmpl):
    p = tmpl.params(n_subunits=2000, **CFG0)
    return _perturbed_state(tmpl, 2000, p.box, 2000)


def test_forces_2000_subunits(tmpl, state2000):
    pos, vel = state2000
    p = tmpl.params(n_subunits=2000, **CFG0)
    s = sg.System(tmpl, 2000, p)
    s.set_state(pos, vel)
    got = s.forces()
    o = Oracle(2000, L=0.3, B=5, **CFG0)
    o.set_state(pos, vel)
    ref = o.forces()
    _check_forces(got, ref)
    assert np.linalg.norm(ref[3]) > 0 and np.linalg.norm(ref[4]) > 0
    # the Yukawa cell grid of this size has more than three cells per axis
    assert p.box / (p.ecut_el + p.skin) > 12
    # FP32 pair arithmetic against the oracle's FP64 forces
    p32 = tmpl.params(n_subunits=2000, **CFG0)
    p32.pair_fp32 = 1
    s32 = sg.System(tmpl, 2000, p32)
    s32.set_state(pos, vel)
    err = rel_norm(s32.forces()[0], ref[0])
    assert 0 < err < 1e-6, err
    o.close()


@pytest.mark.parametrize("brownian", [False, True])
def test_trajectory_2000_subunits(tmpl, state2000, brownian):
    pos, vel = state2000
    p = tmpl.params(n_subunits=2000, brownian=1 if brownian else 0, **CFG0)
    s = sg.System(tmpl, 2000, p)
    s.set_state(pos, vel)
    o = Oracle(2000, L=0.3, B=5, brownian=brownian, **CFG0)
    o.set_state(pos, vel)
    s.step(100)
    o.time_steps(100, 0)
    gp, gv = s.get_state()
    op, ov = o.get_state()[:2]
    assert rel_norm(gp, op) < 1e-8, rel_norm(gp, op)
    assert rel_norm(gv, ov) < 1e-8, rel_norm(gv, ov)
    if not brownian:
        gxi, geta = s.thermostat()
        oxi, oeta = o.get_state()[2:]
        np.testing.assert_allclose(gxi, oxi, rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(geta, oeta, rtol=1e-8, atol=1e-12)
    o.close()


# ---- 256 replicas -------------------------------------------------------------------------

def test_replica_sweep_256(tmpl):
    sizes, params, pos, vel, draws = replica_sweep.sweep(tmpl, 256, 4, fp32=False)
    assert len(sizes) == 256 and min(sizes) >= 20 and max(sizes) <= 400
    batch = sg.System(tmpl, sizes, params)
    batch.set_state(np.concatenate(pos), np.concatenate(vel))
    batch.step(100)
    bp, bv = batch.get_state()
    fb = batch.forces()
    offs = np.concatenate([[0], np.cumsum(sizes)]) * tmpl.n_beads
    salts = [d["c_mM"] for d in draws]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmin(salts)), int(np.argmax(salts))})
    for k in picks:
        sl = slice(offs[k], offs[k + 1])
        single = sg.System(tmpl, sizes[k], params[k])
        single.set_state(pos[k], vel[k])
        single.step(100)
        sp, sv = single.get_state()
        assert np.array_equal(bp[sl], sp) and np.array_equal(bv[sl], sv), k
        d = draws[k]
        L = replica_sweep.reference_list_radius(TEMPLATE, params[k].ecut, params[k].ecut_el)
        o = Oracle(sizes[k], C_uM=800.0, c_mM=d["c_mM"], elj_att=d["elj_att"], K=d["K"], dt=0.005, L=L, B=5)
        o.set_state(sp, sv)
        _check_forces([f[sl] for f in fb], o.forces())
        o.close()
        single.close()


# ---- 5000 subunits, uncut Yukawa ----------------------------------------------------------

def _min_image(d, box):
    # single conditional shift per axis (src/functions.h:42-56)
    return np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))


def test_far_field_5000_subunits(tmpl):
    rf = 20.0
    p = tmpl.params(n_subunits=5000, C_uM=800.0, c_mM=10.0, elj_att=2.1, dt=0.005, K=298.0, ecut_c=0.0)
    p.far_rf = rf
    pos, vel = tmpl.make_state(5000, p.box, p.T, 5000)
    s = sg.System(tmpl, 5000, p)
    s.set_state(pos, vel)
    s.step(50)
    pos, vel = s.get_state()
    f_full = s.forces()[3]
    s.far_drop(True)
    f_near = s.forces()[3]
    s.far_drop(False)
    far_gpu = f_full - f_near
    _, _, q, _ = read_template_xyz(TEMPLATE)
    beads = read_template(TEMPLATE)[0]
    sig = np.array([b[5] for b in beads])
    np_ = tmpl.n_beads
    qi = np.nonzero(q)[0]
    c = tmpl.centre
    centres = pos[c::np_]
    kappa, lb = p.kappa, p.lb
    qpos = pos.reshape(5000, np_, 3)[:, qi]                    # [Ns][nq][3]
    for A in (0, 1777, 4999):
        d = _min_image(centres[A] - centres, p.box)
        far = np.nonzero(np.einsum("ij,ij->i", d, d) > rf * rf)[0]
        assert len(far) > 4000
        F = np.zeros((len(qi), 3))
        for a, ia in enumerate(qi):
            r = _min_image(qpos[A, a][None, None, :] - qpos[far], p.box)   # [nB][nq][3]
            rr = np.sqrt(np.einsum("ijk,ijk->ij", r, r))
            shc = (sig[ia] + sig[qi]) / 2
            Y = np.exp(kappa * shc) / ((1 + kappa * sig[ia] / 2) * (1 + kappa * sig[qi] / 2))
            f = q[ia] * q[qi] * lb * np.exp(-kappa * rr) / rr ** 2 * (kappa + 1 / rr) * Y
            F[a] = np.einsum("ij,ijk->k", f, r)
        got = far_gpu[A * np_ + qi]
        assert np.linalg.norm(F) > 0
        assert rel_norm(got, F) < 1e-10, (A, rel_norm(got, F))
        # beads without charge get no far field
        assert np.all(far_gpu[A * np_:(A + 1) * np_][q == 0] == 0)


# ---- asymmetric pair rules ----------------------------------------------------------------

@pytest.fixture(scope="module")
def rules_template(tmp_path_factory):
    beads, edges, faces, _, ss, sm = read_template(TEMPLATE)
    out = []
    for k, (x, y, z, t, q, d, m) in enumerate(beads):
        if t == 2:
            d, m = 1.2, 1.3
        if k in (7, 30, 62):
            t, d, m = 7, 0.9, 0.8
        out.append((x, y, z, t, q, d, m))
    path = str(tmp_path_factory.mktemp("rules") / "trimer_rules")
    return write_template(path, out, edges, faces, [(1, 2), (2, 7)], ss, sm)


def _rules_state(path, box, seed, spacing):
    xyz = read_template_xyz(path)[0]
    return make_state(xyz, 27, box, seed, spacing=spacing, perturb=0.03, min_gap=1.0)


def test_pair_rules_forces_and_energies(rules_template):
    t = sg.Template(rules_template)
    p = t.params(n_subunits=27, **CFG0)
    pos, vel = _rules_state(rules_template, p.box, 51, 7.7)
    s = sg.System(t, 27, p)
    s.set_state(pos, vel)
    got = s.forces()
    o = Oracle(27, template=rules_template, L=3.0, B=5, **CFG0)
    o.set_state(pos, vel)
    ref = o.forces()
    _check_forces(got, ref)
    # both ordered rows have pairs inside the attractive cutoff in this state
    types = np.array([b[3] for b in read_template(rules_template)[0]] * 27)
    sub = np.repeat(np.arange(27), t.n_beads)
    d = pos[:, None, :] - pos[None, :, :]
    d = np.where(d > p.box / 2, d - p.box, np.where(d < -p.box / 2, d + p.box, d))
    close = (np.einsum("ijk,ijk->ij", d, d) < p.ecut ** 2) & (sub[:, None] != sub[None, :])
    assert (close & (types[:, None] == 1) & (types[None, :] == 2)).any()
    assert (close & (types[:, None] == 2) & (types[None, :] == 7)).any()
    # the attractive and WCA terms both act, and Newton's third law does not hold for them
    assert np.linalg.norm(ref[4]) > 0
    net = np.abs(ref[0].sum(axis=0)).max()
    assert net > 1e-6 * np.abs(ref[0]).max(), net
    ge, gpb = s.energies(per_bead=True)
    oe, opb = o.energies(per_bead=True)
    for k in range(5):
        assert abs(ge[0][k] - oe[k]) <= 1e-10 * max(abs(oe[k]), 1e-300), (k, ge[0][k], oe[k])
        assert rel_norm(gpb[:, k], opb[:, k]) < 1e-10, k
    p32 = t.params(n_subunits=27, **CFG0)
    p32.pair_fp32 = 1
    s32 = sg.System(t, 27, p32)
    s32.set_state(pos, vel)
    err = rel_norm(s32.forces()[0], ref[0])
    assert 0 < err < 1e-6, err


@pytest.mark.parametrize("brownian", [False, True])
def test_pair_rules_trajectory(rules_template, brownian):
    t = sg.Template(rules_template)
    p = t.params(n_subunits=27, brownian=1 if brownian else 0, **CFG0)
    pos, vel = _rules_state(rules_template, p.box, 52, 7.7)
    s = sg.System(t, 27, p)
    s.set_state(pos, vel)
    o = Oracle(27, template=rules_template, L=3.0, B=5, brownian=brownian, **CFG0)
    o.set_state(pos, vel)
    s.step(200)
    o.run(200)
    gp, gv = s.get_state()
    op, ov = o.get_state()[:2]
    assert rel_norm(gp, op) < 1e-8, rel_norm(gp, op)
    assert rel_norm(gv, ov) < 1e-8, rel_norm(gv, ov)
    # the net force is not zero, so the centre-of-mass velocity removal acts; both remove it
    assert s.stats()["vcom"] > 0
    assert np.abs(gv.sum(axis=0)).max() < 1e-9 * np.abs(gv).max() * len(gv)
