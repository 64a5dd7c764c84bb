"""Smallest systems against the CPU oracle: one subunit (no LJ list at all: bonded terms and the
Yukawa terms among its own charged beads) and two subunits in a dilute box, forces, energies and
NH / BD trajectories.

Tolerances as in test_gpu_parity: forces and energies 1e-10 norm-wise relative, positions and
velocities after 100 steps 1e-8.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

CFG = dict(C_uM=20.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)


def _pair(ns, brownian=False, seed=3):
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=ns, brownian=1 if brownian else 0, **CFG)
    p.skin = 0.3
    xyz = read_template_xyz(TEMPLATE)[0]
    pos, vel = make_state(xyz, ns, p.box, seed, spacing=8.4 if ns > 1 else None)
    o = Oracle(ns, brownian=brownian, L=3.0, B=5, **CFG)
    o.set_state(pos, vel)
    s = sg.System(t, ns, p)
    s.set_state(pos, vel)
    return s, o, pos, vel


@pytest.mark.parametrize("ns", [1, 2])
def test_small_forces_and_energies(ns):
    s, o, _, _ = _pair(ns)
    ref = o.forces()
    got = s.forces()
    for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
        if np.linalg.norm(r) == 0:
            assert np.linalg.norm(g) == 0, name
            continue
        assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    if ns == 1:  # no LJ pairs within a subunit; its charged beads interact
        assert np.linalg.norm(got[4]) == 0 and s.stats()["lj_entries"] == 0
        assert np.linalg.norm(got[3]) > 0 and s.stats()["q_entries"] > 0
    ge = s.energies()[0]
    oe = o.energies()
    np.testing.assert_allclose(ge, oe, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("ns,brownian", [(1, False), (1, True), (2, False), (2, True)])
def test_small_trajectories(ns, brownian):
    s, o, _, _ = _pair(ns, brownian=brownian, seed=5)
    n = 100
    o.run(n)
    s.step(n)
    op, ov, _, _ = o.get_state()
    gp, gv = s.get_state()
    assert rel_norm(gp, op) < 1e-8
    assert rel_norm(gv, ov) < 1e-8
    assert s.stats()["steps"] == n
