"""GPU path (through the C ABI) against the CPU oracle on identical seeded states.

Tolerances:
  forces and energies: 1e-10 norm-wise relative (the reference's own arithmetic is x87 long
  double; the FP64 GPU path and the FP64 oracle differ by summation order only);
  trajectories: positions/velocities after 100-200 steps within 1e-8 (chaotic growth of
  round-off stays far below this over one reduced time unit);
  BD noise: bit-exact against libstdc++'s std::default_random_engine stream.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, TEMPLATE_BIN, Oracle, bd_uniforms, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)


def _system(ns, template=TEMPLATE, skin=0.3, fp32=False, brownian=False, q_chain=5, **over):
    t = sg.Template(template)
    cli = dict(CFG0, n_subunits=ns, brownian=1 if brownian else 0, q_chain=q_chain)
    cli.update(over)
    p = t.params(**cli)
    p.skin = skin
    p.pair_fp32 = 1 if fp32 else 0
    return t, p, sg.System(t, ns, p)


def _state(template, ns, box, seed, spacing, perturb=0.05):
    xyz = read_template_xyz(template)[0]
    return make_state(xyz, ns, box, seed, spacing=spacing, perturb=perturb)


def _oracle(ns, template=TEMPLATE, brownian=False, q_chain=5, long_double=False, **over):
    kw = dict(CFG0)
    kw.update(over)
    return Oracle(ns, template=template, brownian=brownian, q_chain=q_chain, long_double=long_double, **kw)


@pytest.mark.parametrize("template,seed,spacing", [(TEMPLATE, 1, 8.6), (TEMPLATE, 2, 8.3), (TEMPLATE_BIN, 3, 8.2)])
def test_forces_match_oracle(template, seed, spacing):
    t, p, sysg = _system(20, template)
    pos, vel = _state(template, 20, p.box, seed, spacing)
    o = _oracle(20, template)
    o.set_state(pos, vel)
    ref = o.forces()
    sysg.set_state(pos, vel)
    got = sysg.forces()
    names = ["total", "stretch", "bend", "yukawa", "lj"]
    for name, g, r in zip(names, got, ref):
        if np.linalg.norm(r) == 0:
            assert np.linalg.norm(g) == 0, name
            continue
        assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    assert np.linalg.norm(ref[4]) > 0, "state has no LJ contacts"
    if template == TEMPLATE:
        assert np.linalg.norm(ref[3]) > 0


def test_forces_under_translation_and_outside_the_box():
    """One periodic state in three representations: as placed (subunits around the origin), moved
    by half a box along x and wrapped into the box (subunits straddling the x faces), and moved
    without wrapping (coordinates up to ~14 sigma outside the box). Forces of all three agree with
    the oracle's at 1e-10. The force kernel takes plain coordinate differences only where they equal
    the minimum image (a subunit inside a quarter box of its centre bead; beads away from the faces
    when every position was given inside the box), so these cover its shifted and unshifted paths."""
    t, p, sysg = _system(20)
    pos, vel = _state(TEMPLATE, 20, p.box, 11, 8.3)
    h = p.box / 2
    o = _oracle(20)
    o.set_state(pos, vel)
    ref = o.forces()
    moved = pos + np.array([h, 0.0, 0.0])
    wrapped = np.where(moved > h, moved - p.box, moved)
    assert (np.abs(moved) > h).sum() > 100
    assert (wrapped[:, 0] > h - 2).any() and (wrapped[:, 0] < -h + 2).any()
    for x in (pos, wrapped, moved):
        sysg.set_state(x, vel)
        got = sysg.forces()
        for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
            assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    assert np.linalg.norm(ref[4]) > 0 and np.linalg.norm(ref[3]) > 0


@pytest.mark.parametrize("template,seed,spacing", [(TEMPLATE, 4, 8.6), (TEMPLATE_BIN, 3, 8.2)])
def test_energies_match_oracle(template, seed, spacing):
    """Energy totals and per-bead terms; both copies of the template (the bin/infiles one has
    sigma = 0.8889 nm and no charges: its Yukawa terms are exactly zero on both sides)."""
    t, p, sysg = _system(20, template)
    pos, vel = _state(template, 20, p.box, seed, spacing)
    o = _oracle(20, template)
    o.set_state(pos, vel)
    ref, ref_pb = o.energies(per_bead=True)
    sysg.set_state(pos, vel)
    got, got_pb = sysg.energies(per_bead=True)
    got = got[0]
    for k, name in enumerate(["ke", "se", "be", "ne", "ce"]):
        assert abs(got[k] - ref[k]) <= 1e-10 * max(abs(ref[k]), 1e-300), (name, got[k], ref[k])
    # per bead: se, be, ne, ce, ke in the reference's BEAD fields
    for k in range(5):
        if np.linalg.norm(ref_pb[:, k]) == 0:
            assert np.linalg.norm(got_pb[:, k]) == 0, k
            continue
        assert rel_norm(got_pb[:, k], ref_pb[:, k]) < 1e-10, k
    assert ref[3] != 0, "state has no LJ contacts"
    if template == TEMPLATE_BIN:
        assert ref[4] == 0 and got[4] == 0


@pytest.mark.parametrize("q_chain,template,spacing", [(5, TEMPLATE, 8.6), (1, TEMPLATE, 8.6), (5, TEMPLATE_BIN, 8.2)])
def test_md_trajectory_matches_oracle(q_chain, template, spacing):
    t, p, sysg = _system(20, template, q_chain=q_chain)
    pos, vel = _state(template, 20, p.box, 5, spacing)
    o = _oracle(20, template, q_chain=q_chain, L=3.0, B=5)
    o.set_state(pos, vel)
    sysg.set_state(pos, vel)
    n = 200
    o.run(n)
    sysg.step(n)
    op, ov, oxi, oeta = o.get_state()
    gp, gv = sysg.get_state()
    assert rel_norm(gp, op) < 1e-8
    assert rel_norm(gv, ov) < 1e-8
    gxi, geta = sysg.thermostat()
    np.testing.assert_allclose(gxi, oxi, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(geta, oeta, rtol=1e-8, atol=1e-12)
    ge = sysg.energies()[0]
    oe = o.energies()
    np.testing.assert_allclose(ge, oe, rtol=1e-8, atol=1e-9)
    assert sysg.stats()["steps"] == n


def test_bd_noise_stream_bit_exact():
    t, p, sysg = _system(20, brownian=True)
    nb = 20 * t.n_beads
    for step in [0, 1, 7]:
        got = sysg.bd_uniforms(0, step).reshape(-1)
        ref = bd_uniforms(step * 3 * nb, 3 * nb)
        assert np.array_equal(got, ref), step


@pytest.mark.parametrize("template,spacing", [(TEMPLATE, 8.6), (TEMPLATE_BIN, 8.2)])
def test_bd_trajectory_matches_oracle(template, spacing):
    t, p, sysg = _system(20, template, brownian=True)
    pos, vel = _state(template, 20, p.box, 6, spacing)
    o = _oracle(20, template, brownian=True, L=3.0, B=5)
    o.set_state(pos, vel)
    sysg.set_state(pos, vel)
    n = 100
    o.run(n)
    sysg.step(n)
    op, ov, _, _ = o.get_state()
    gp, gv = sysg.get_state()
    assert rel_norm(gp, op) < 1e-8
    assert rel_norm(gv, ov) < 1e-8


def test_deterministic_rerun():
    t, p, sysg = _system(20)
    pos, vel = _state(TEMPLATE, 20, p.box, 5, 8.6)
    outs = []
    for _ in range(2):
        sysg.set_state(pos, vel)
        sysg.step(150)
        outs.append(sysg.get_state())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_fp32_pair_form_error():
    """FP32 pair arithmetic against the oracle's FP64 forces (measured ~1e-8 on this state)."""
    _, p32, s32 = _system(20, fp32=True)
    pos, vel = _state(TEMPLATE, 20, p32.box, 9, 8.6)
    o = _oracle(20)
    o.set_state(pos, vel)
    ref = o.forces()
    s32.set_state(pos, vel)
    got = s32.forces()
    assert 0 < rel_norm(got[0], ref[0]) < 1e-6, rel_norm(got[0], ref[0])
    # bonded terms stay FP64
    assert rel_norm(got[1], ref[1]) < 1e-10 and rel_norm(got[2], ref[2]) < 1e-10


def test_replicas_match_single_runs():
    t = sg.Template(TEMPLATE)
    sizes = [20, 27, 12]
    params = []
    for k, ns in enumerate(sizes):
        p = t.params(C_uM=800.0, c_mM=[100.0, 300.0, 150.0][k], elj_att=[2.1, 1.6, 2.4][k], dt=0.005,
                     K=[298.0, 285.0, 310.0][k], n_subunits=ns)
        params.append(p)
    states = [_state(TEMPLATE, ns, params[k].box, 20 + k, 8.7) for k, ns in enumerate(sizes)]
    batch = sg.System(t, sizes, params)
    batch.set_state(np.concatenate([s[0] for s in states]), np.concatenate([s[1] for s in states]))
    batch.step(120)
    bp, bv = batch.get_state()
    off = 0
    for k, ns in enumerate(sizes):
        single = sg.System(t, ns, params[k])
        single.set_state(*states[k])
        single.step(120)
        sp, sv = single.get_state()
        n = ns * t.n_beads
        assert np.array_equal(bp[off:off + n], sp), k
        assert np.array_equal(bv[off:off + n], sv), k
        off += n


@pytest.mark.parametrize("brownian", [False, True])
def test_list_regrowth_during_steps(brownian):
    """A dense state outgrows the initial list capacities inside sf_step: the chunk is replayed
    with larger lists (and a re-captured graph) and the trajectory still follows the oracle."""
    t, p, sysg = _system(27, brownian=brownian)
    pos, vel = _state(TEMPLATE, 27, p.box, 11, 7.6, perturb=0.02)
    o = _oracle(27, brownian=brownian, L=3.0, B=5)
    o.set_state(pos, vel)
    sysg.set_state(pos, vel)
    caps0 = (sysg.stats()["lj_capacity"], sysg.stats()["q_capacity"])
    o.run(150)
    sysg.step(150)
    st = sysg.stats()
    op, ov = o.get_state()[:2]
    gp, gv = sysg.get_state()
    assert rel_norm(gp, op) < 1e-8 and rel_norm(gv, ov) < 1e-8
    assert st["steps"] == 150
    assert caps0[0] >= 24 and caps0[1] >= 48


def test_dense_state_forces():
    t, p, sysg = _system(27)
    pos, vel = _state(TEMPLATE, 27, p.box, 11, 7.6, perturb=0.02)
    o = _oracle(27)
    o.set_state(pos, vel)
    ref = o.forces()[0]
    sysg.set_state(pos, vel)
    got = sysg.forces()[0]
    assert rel_norm(got, ref) < 1e-10
