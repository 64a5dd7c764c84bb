"""Uncut Yukawa (the reference's -e 0 with a list radius covering the box: every charged pair
interacts) on the GPU: near field from the lists for subunit pairs whose centre beads are
within far_rf, far field subunit pair by subunit pair. Against the oracle, which sums every
pair of its all-covering list.

Tolerances as in test_gpu_parity: forces and energies 1e-10 norm-wise, trajectories 1e-8.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, elj_att=2.1, dt=0.005, K=298.0, ecut_c=0.0)


def _setup(c_mM, far_rf, ns=20, seed=1, spacing=8.6, fp32=False):
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=ns, c_mM=c_mM, **CFG0)
    p.far_rf = far_rf
    p.pair_fp32 = 1 if fp32 else 0
    xyz = read_template_xyz(TEMPLATE)[0]
    pos, vel = make_state(xyz, ns, p.box, seed, spacing=spacing, perturb=0.05)
    s = sg.System(t, ns, p)
    s.set_state(pos, vel)
    o = Oracle(ns, template=TEMPLATE, c_mM=c_mM, L=2.0 * p.box, B=1, **CFG0)
    o.set_state(pos, vel)
    return t, p, s, o, pos, vel


@pytest.mark.parametrize("c_mM,far_rf", [(10.0, 10.0), (100.0, 10.0), (10.0, 6.0)])
def test_uncut_forces_and_energies(c_mM, far_rf):
    t, p, s, o, pos, vel = _setup(c_mM, far_rf)
    assert p.ecut_el == 0.0
    ref = o.forces()
    got = s.forces()
    for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
        assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    # the far field is a real part of the Yukawa force here
    assert np.linalg.norm(ref[3]) > 0
    ref_e, ref_pb = o.energies(per_bead=True)
    got_e, got_pb = s.energies(per_bead=True)
    for k in range(5):
        assert abs(got_e[0][k] - ref_e[k]) <= 1e-10 * max(abs(ref_e[k]), 1e-300), k
        assert rel_norm(got_pb[:, k], ref_pb[:, k]) < 1e-10, k


def test_uncut_split_radius_does_not_change_forces():
    """The split radius only moves pairs between the two sums."""
    _, _, s1, _, pos, vel = _setup(10.0, 6.0)
    _, _, s2, _, _, _ = _setup(10.0, 14.0)
    f1 = s1.forces()[3]
    f2 = s2.forces()[3]
    assert rel_norm(f1, f2) < 1e-13


def test_uncut_trajectory_matches_oracle():
    t, p, s, o, pos, vel = _setup(10.0, 10.0)
    o.run(150)
    s.step(150)
    op, ov = o.get_state()[:2]
    gp, gv = s.get_state()
    assert rel_norm(gp, op) < 1e-8 and rel_norm(gv, ov) < 1e-8
