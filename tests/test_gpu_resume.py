"""State transfer through the C ABI beyond the reference's restart:

- sf_get_resume + sf_set_state_resume continue a run bit for bit in another handle (NH MD and BD):
  100 steps, transfer, 100 steps equal 200 uninterrupted steps exactly;
- sf_set_positions + sf_forces_ex (the forceCalculation drop-in for callers with their own
  integrator) give the forces of sf_forces bit for bit, keep velocities and the thermostat, and
  agree with the oracle at 1e-10;
- a replica whose velocities become non-finite stops sf_step with SF_E_NONFINITE at the last
  finite state. The condition is set up from the host: two one-bead subunits (sigma 0.01, no
  charge, no attraction) move head-on at constant velocity under NVE (q = 1) so that they sit on
  the same point after exactly 10 steps; the WCA term of a pair at distance 0 is 0 x inf in
  ordinary floating point (no fault), the velocities become NaN in step 10, and the handle must
  be left after step 9 with the beads at -0.5 and +0.5.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz
from template_writer import write_template

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)


def _system(t, ns, brownian=False):
    p = t.params(n_subunits=ns, brownian=1 if brownian else 0, **CFG0)
    return sg.System(t, ns, p), p


@pytest.mark.parametrize("brownian", [False, True])
def test_resume_is_bit_continuous(brownian):
    t = sg.Template(TEMPLATE)
    a, p = _system(t, 20, brownian)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 20, p.box, 31, spacing=8.6)
    a.set_state(pos, vel)
    a.step(100)
    mid_pos, mid_vel = a.get_state()
    recs, tforce = a.get_resume()
    assert recs[0]["step"] == 100
    a.step(100)
    b, _ = _system(t, 20, brownian)
    b.set_state_resume(mid_pos, mid_vel, recs, tforce)
    b.step(100)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    ra, _ = a.get_resume(with_force=False)
    rb, _ = b.get_resume(with_force=False)
    assert ra == rb
    # an uninterrupted 200-step run from the same start
    c, _ = _system(t, 20, brownian)
    c.set_state(pos, vel)
    c.step(200)
    pc, vc = c.get_state()
    assert np.array_equal(pa, pc) and np.array_equal(va, vc)
    # the reference's restart (thermostat and engine afresh) is a different continuation
    d, _ = _system(t, 20, brownian)
    d.set_state(mid_pos, mid_vel)
    d.step(100)
    assert not np.array_equal(d.get_state()[1], va)


def test_resume_record_continues_in_the_oracle():
    """The resume record means what the reference's thermostat state means: the oracle, given
    positions, velocities and the record's xi, eta and KE, continues the GPU run at 1e-8."""
    t = sg.Template(TEMPLATE)
    a, p = _system(t, 20)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 20, p.box, 34, spacing=8.6)
    a.set_state(pos, vel)
    a.step(100)
    mid_pos, mid_vel = a.get_state()
    rec = a.get_resume(with_force=False)[0][0]
    a.step(100)
    ga, gv = a.get_state()
    gxi, _ = a.thermostat()
    o = Oracle(20, B=1, **CFG0)
    o.set_state(mid_pos, mid_vel)
    o.set_thermostat(rec["xi"], rec["eta"], rec["ke"], step_index=100)
    o.run(100)
    op, ov, oxi, _ = o.get_state()
    assert rel_norm(gv, ov) < 1e-8 and rel_norm(ga, op) < 1e-8
    assert np.allclose(gxi[:4], oxi[:4], rtol=1e-8, atol=1e-12)


def test_resume_checks_its_records():
    t = sg.Template(TEMPLATE)
    a, p = _system(t, 20, True)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 20, p.box, 32, spacing=8.6)
    a.set_state(pos, vel)
    recs, tforce = a.get_resume()
    with pytest.raises(sg.SfError):  # BD without the integrator force
        a.set_state_resume(pos, vel, recs, None)
    bad = [dict(recs[0], rng=0)]
    with pytest.raises(sg.SfError):
        a.set_state_resume(pos, vel, bad, tforce)


def test_set_positions_forces_drop_in():
    t = sg.Template(TEMPLATE)
    s, p = _system(t, 20)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 20, p.box, 33, spacing=8.6)
    s.set_state(pos, vel)
    f0 = s.forces()
    s.step(60)
    pos1, vel1 = s.get_state()
    xi1, eta1 = s.thermostat()
    # back to the start positions: lists are rebuilt (the beads moved) and the forces are the
    # start's bit for bit; velocities and the chain are untouched
    s.set_positions(pos)
    f1 = s.forces_ex(False)
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)
    assert np.array_equal(s.get_state()[1], vel1)
    xi2, eta2 = s.thermostat()
    assert np.array_equal(xi1, xi2) and np.array_equal(eta1, eta2)
    # a displacement below skin/2 keeps the lists: the same forces as with a fresh build
    rng = np.random.default_rng(4)
    moved = pos1 + rng.uniform(-0.02, 0.02, pos1.shape)
    s.set_positions(moved)
    fa = s.forces_ex(False)
    fb = s.forces_ex(True)
    for a, b in zip(fa, fb):
        assert np.array_equal(a, b)
    o = Oracle(20, **CFG0)
    o.set_state(moved, vel1)
    ref = o.forces()
    for g, r in zip(fa, ref):
        if np.linalg.norm(r) > 0:
            assert rel_norm(g, r) < 1e-10


def test_nonfinite_state_stops_at_last_finite_step(tmp_path):
    path = write_template(str(tmp_path / "point"), beads=[(0.0, 0.0, 0.0, 1, 0.0, 0.01, 1.0)])
    t = sg.Template(path)
    p = t.params(n_subunits=2, C_uM=50.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, q_chain=1)
    assert p.box / 2 > 5.5
    s = sg.System(t, 2, p)
    pos = np.array([[-5.0, 0.0, 0.0], [5.0, 0.0, 0.0]])
    vel = np.array([[100.0, 0.0, 0.0], [-100.0, 0.0, 0.0]])
    s.set_state(pos, vel)
    s.step(5)  # finite: each bead moves 0.5 sigma per step exactly
    gp, gv = s.get_state()
    assert np.array_equal(gp[:, 0], [-2.5, 2.5]) and np.array_equal(gv, vel)
    with pytest.raises(sg.SfError) as e:
        s.step(50)
    assert e.value.code == sg.SF_E_NONFINITE
    assert "step 10" in str(e.value) and "step 9" in str(e.value), str(e.value)
    gp, gv = s.get_state()
    assert np.array_equal(gp[:, 0], [-0.5, 0.5]) and np.all(gp[:, 1:] == 0)
    assert np.array_equal(gv, vel)
    assert s.stats()["steps"] == 9
    assert s.get_resume(with_force=False)[0][0]["step"] == 9
