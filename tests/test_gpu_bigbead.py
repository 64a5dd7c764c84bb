"""Encapsulation big beads on the GPU (through the C ABI) against the CPU oracle.

Reference: forceCalculation_bigbead src/forces.cpp:174-236, update_LJ_ES_energies_bigbeads
src/energies.cpp:96-136, the big beads' BD draws src/md.cpp:482-487, placement
src/initialize.cpp:566-680, defaults src/md.cpp:86-90 (diameter 6, charge -78, shc 3.5).

Tolerances as for the other parity tests: forces and energies 1e-10 norm-wise relative to the
long double oracle (the reference's arithmetic), trajectories 1e-8 after 100 steps against the
FP64 oracle; the BD noise stream bit-exact.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, bd_uniforms, place_big_beads, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
BIG = dict(sigma=6.0, charge=-78.0, shc=3.5)


def _system(ns, brownian=False):
    t = sg.Template(TEMPLATE)
    p = t.params(**dict(CFG0, n_subunits=ns, brownian=1 if brownian else 0, q_chain=5))
    p.skin = 0.3
    return t, p, sg.System(t, ns, p)


def _oracle(ns, brownian=False, long_double=False, **over):
    kw = dict(CFG0)
    kw.update(over)
    return Oracle(ns, brownian=brownian, long_double=long_double, **kw)


def _setup(seed, box, o):
    """State and two big beads: one placed by the reference's rule, then moved to 3.7 from its
    nearest bead (inside the shifted-WCA range), and a second placed bead."""
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 20, box, seed, spacing=8.6, perturb=0.05)
    sig, _ = o.bead_props()
    c = place_big_beads(pos, sig, box, 2, BIG["sigma"])
    d = pos - c[0]
    d -= box * np.floor(d / box + 0.5)
    r = np.linalg.norm(d, axis=1)
    j = int(np.argmin(r))
    c[0] = c[0] + d[j] * (1.0 - 3.7 / r[j])
    c[0] -= box * np.floor(c[0] / box + 0.5)
    return pos, vel, c


def test_big_bead_forces_and_energies_match_oracle():
    t, p, sysg = _system(20)
    o = _oracle(20, long_double=True)
    pos, vel, c = _setup(11, p.box, o)
    o.set_big_beads(c, BIG["sigma"], BIG["charge"], BIG["shc"])
    o.set_state(pos, vel)
    ref = o.forces()[0]
    ref_e, ref_pb = o.energies(per_bead=True)
    ref_big = o.big_energies()
    sysg.set_big_beads(c, **BIG)
    sysg.set_state(pos, vel)
    got = sysg.forces()[0]
    got_e, got_pb = sysg.energies(per_bead=True)
    got_big = sysg.big_energies()
    assert rel_norm(got, ref) < 1e-10, rel_norm(got, ref)
    for k, name in enumerate(["ke", "se", "be", "ne", "ce"]):
        assert abs(got_e[0][k] - ref_e[k]) <= 1e-10 * max(abs(ref_e[k]), 1e-300), (name, got_e[0][k], ref_e[k])
    for k in range(5):
        assert rel_norm(got_pb[:, k], ref_pb[:, k]) < 1e-10, k
    np.testing.assert_allclose(got_big, ref_big, rtol=1e-10)
    assert ref_big[0] != 0 and ref_big[1] > 0
    # the big beads change the forces (otherwise the test proves nothing)
    o.set_big_beads(np.zeros((0, 3)), BIG["sigma"], BIG["charge"], BIG["shc"])
    assert rel_norm(o.forces()[0], ref) > 1e-6


@pytest.mark.parametrize("brownian", [False, True])
def test_big_bead_trajectory_matches_oracle(brownian):
    t, p, sysg = _system(20, brownian=brownian)
    o = _oracle(20, brownian=brownian, L=3.0, B=5)
    pos, vel, c = _setup(12, p.box, o)
    o.set_big_beads(c, BIG["sigma"], BIG["charge"], BIG["shc"])
    o.set_state(pos, vel)
    sysg.set_big_beads(c, **BIG)
    sysg.set_state(pos, vel)
    n = 100
    o.run(n)
    sysg.step(n)
    op, ov, _, _ = o.get_state()
    gp, gv = sysg.get_state()
    assert rel_norm(gp, op) < 1e-8, rel_norm(gp, op)
    assert rel_norm(gv, ov) < 1e-8, rel_norm(gv, ov)


def test_bd_noise_stream_skips_big_bead_draws():
    t, p, sysg = _system(20, brownian=True)
    nb = 20 * t.n_beads
    sysg.set_big_beads(np.zeros((3, 3)), **BIG)
    for step in [0, 1, 7]:
        got = sysg.bd_uniforms(0, step).reshape(-1)
        assert np.array_equal(got, bd_uniforms(step * 3 * (nb + 3), 3 * nb)), step


def test_step_needs_restart_after_big_bead_change():
    t, p, sysg = _system(20)
    o = _oracle(20)
    pos, vel, c = _setup(13, p.box, o)
    sysg.set_state(pos, vel)
    sysg.set_big_beads(c, **BIG)
    with pytest.raises(sg.SfError):
        sysg.step(1)
    sysg.set_state(pos, vel)
    sysg.step(1)


def test_removing_big_beads_restores_the_plain_path():
    """n = 0 removes the big beads: the run is then bitwise the run of a handle that never had
    them (forces, noise stream multiplier and graph all back to the plain path)."""
    outs = []
    for with_then_without in (False, True):
        t, p, sysg = _system(20, brownian=True)
        o = _oracle(20)
        pos, vel, c = _setup(14, p.box, o)
        if with_then_without:
            sysg.set_big_beads(c, **BIG)
            sysg.set_state(pos, vel)
            sysg.step(20)
            sysg.set_big_beads(np.zeros((0, 3)), **BIG)
        sysg.set_state(pos, vel)
        sysg.step(50)
        outs.append(sysg.get_state())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_big_beads_need_a_single_replica_handle():
    t = sg.Template(TEMPLATE)
    p = t.params(**dict(CFG0, n_subunits=20, q_chain=5))
    batch = sg.System(t, [20, 20], [p, p])
    with pytest.raises(sg.SfError):
        batch.set_big_beads(np.zeros((1, 3)), **BIG)
