"""Physical parameters across the ranges the reference's CLI takes (salt -c, attraction -E,
temperature -K, chain length -q, concentration -C, spring constants -s / -k, Q, BD damping -d)
against the oracle, 20 subunits: forces at 1e-10 and 100-step trajectories at 1e-8 (as in
test_gpu_parity). One case per row: a fractional design over the parameters, not a full grid.
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz

pytestmark = pytest.mark.gpu

CASES = [
    # c_mM, elj_att, K, q_chain, C_uM, ks, kb, Q, brownian, damp
    (50.0, 1.5, 280.0, 1, 800.0, 50.0, 20.0, 10.0, False, 100.0),
    (800.0, 2.5, 320.0, 3, 800.0, 50.0, 20.0, 10.0, False, 100.0),
    (200.0, 1.8, 300.0, 5, 400.0, 80.0, 10.0, 5.0, False, 100.0),
    (400.0, 2.2, 290.0, 2, 1200.0, 30.0, 40.0, 20.0, False, 100.0),
    (100.0, 2.1, 298.0, 5, 800.0, 50.0, 20.0, 10.0, True, 10.0),
    (600.0, 1.6, 310.0, 5, 600.0, 50.0, 30.0, 10.0, True, 400.0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"c{c[0]:g}-E{c[1]:g}-K{c[2]:g}-q{c[3]}-{'bd' if c[8] else 'nh'}" for c in CASES])
def test_parameters_against_oracle(case):
    c_mM, eps, K, q, C_uM, ks, kb, Q, brownian, damp = case
    kw = dict(C_uM=C_uM, c_mM=c_mM, elj_att=eps, dt=0.005, K=K, ks=ks, kb=kb, Q=Q, damp=damp)
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=20, q_chain=q, brownian=1 if brownian else 0, **kw)
    p.skin = 0.3
    xyz = read_template_xyz(TEMPLATE)[0]
    pos, vel = make_state(xyz, 20, p.box, 17, spacing=8.5 * (800.0 / C_uM) ** (1.0 / 3.0) if C_uM < 800 else 8.5)
    o = Oracle(20, q_chain=q, brownian=brownian, L=3.0, B=5, **kw)
    assert abs(o.box - p.box) < 1e-9 * p.box
    o.set_state(pos, vel)
    s = sg.System(t, 20, p)
    s.set_state(pos, vel)
    ref = o.forces()
    got = s.forces()
    for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
        if np.linalg.norm(r) == 0:
            assert np.linalg.norm(g) == 0, name
            continue
        assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    n = 100
    o.run(n)
    s.step(n)
    op, ov, oxi, oeta = o.get_state()
    gp, gv = s.get_state()
    assert rel_norm(gp, op) < 1e-8
    assert rel_norm(gv, ov) < 1e-8
    if not brownian:
        gxi, geta = s.thermostat()
        np.testing.assert_allclose(gxi, oxi, rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(geta, oeta, rtol=1e-8, atol=1e-12)
