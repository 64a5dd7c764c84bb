"""A dense state that takes the list query's fallback paths, against the oracle: 64 subunits at
4 mM capsomers (box 29.8 sigma) packed 7 sigma apart. Charged beads there have more Yukawa
partners within the list radius than a charged bead's LDS slots hold (128), so their lists are
built from the global cells and sorted in place; the other subunits' records around most
subunits (370-1290, median ~700) exceed the list query's LDS record space (512), so those
subunits take its direct path and the rest the LDS path. Forces and energies at 1e-10, a 50-step NH trajectory at 1e-8 (as in
test_gpu_parity). The state is denser than the reference's own fixed list capacity,
ceil((cutoff + 0.5)^3) per bead (src/md.cpp:231), allows, so the oracle runs with twice it.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

KW = dict(C_uM=4000.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
NS = 64


def _setup():
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=NS, **KW)
    p.skin = 0.3
    xyz, _, q, _ = read_template_xyz(TEMPLATE)
    pos, vel = make_state(xyz, NS, p.box, 3, spacing=7.0)
    return t, p, pos, vel, q


def test_dense_state_exercises_the_overflow_paths():
    _, p, pos, _, q = _setup()
    box = p.box
    qi = np.where(np.tile(q != 0, NS))[0]
    d = pos[qi][:, None, :] - pos[qi][None, :, :]
    d -= box * np.rint(d / box)
    r = np.sqrt((d * d).sum(-1))
    partners = ((r < p.ecut_el + p.skin) & (qi[:, None] != qi[None, :])).sum(1)
    assert partners.max() > 128, partners.max()


def test_dense_forces_energies_trajectory():
    t, p, pos, vel, _ = _setup()
    o = Oracle(NS, L=3.0, B=5, **KW)
    o.set_list_capacity(2 * o.nlv)  # denser than the reference's list capacity is sized for
    o.set_state(pos, vel)
    s = sg.System(t, NS, p)
    s.set_state(pos, vel)
    ref = o.forces()
    got = s.forces()
    for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
        assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    np.testing.assert_allclose(s.energies()[0], o.energies(), rtol=1e-10, atol=1e-12)
    n = 50
    o.run(n)
    s.step(n)
    op, ov, _, _ = o.get_state()
    gp, gv = s.get_state()
    assert rel_norm(gp, op) < 1e-8
    assert rel_norm(gv, ov) < 1e-8
