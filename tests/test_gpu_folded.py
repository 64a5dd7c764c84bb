"""Subunits with bending hinges folded past 90 degrees, GPU against the oracle.

The reference's hinge force (src/edge.cpp:69-188: the face flip makes the product of the two area
vectors non-negative, while its gradient formulas differentiate the unflipped product) is minus
the gradient of Kb (1 - |n0.n1|) only up to a 90 degree fold; past it, it is another function of
the geometry (tests/test_oracle.py::test_bending_folded_hinge pins the oracle there). Such hinges
appear when a subunit is crushed in a cluster or beads of one subunit cross each other (the
model has no excluded volume inside a subunit). Here beads of three subunits are pushed
1.2-1.5 sigma towards their subunit's centre, which folds a few hinges past 90 degrees (asserted),
and forces, energies and a 100-step Nose-Hoover trajectory are compared with the oracle at the
usual tolerances (1e-10 for forces and energies, 1e-8 for the trajectory).
"""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE, Oracle, rel_norm
from souffle_states import make_state, read_template_xyz
from template_writer import read_template

pytestmark = pytest.mark.gpu

CFG0 = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
PUSH = ((0, 94, 1.2), (7, 5, 1.5), (13, 30, 1.2))  # (subunit, template bead, distance in sigma)


def _hinges():
    """Bending hinges (src/edge.cpp: edges of nonzero type with two faces): (x0, x1, across F0,
    across F1) as template bead indices."""
    _, edges, faces, _, _, _ = read_template(TEMPLATE)
    out = []
    for a, b, typ, _ in edges:
        if typ == 0:
            continue
        across = [next(x for x in f if x not in (a, b)) for f in faces if a in f and b in f]
        if len(across) == 2:
            out.append((a, b, across[0], across[1]))
    return out


def _fold_cosines(x, hinges):
    """cos of the fold of each hinge (1 flat, < 0 past 90 degrees) of one subunit's beads x."""
    out = []
    for a, b, c0, c1 in hinges:
        e = x[b] - x[a]
        u, v = np.cross(e, x[c0] - x[a]), np.cross(e, x[c1] - x[a])
        out.append(-np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))
    return np.array(out)


def _folded_state(box, seed):
    xyz = read_template_xyz(TEMPLATE)[0]
    pos, vel = make_state(xyz, 20, box, seed, spacing=8.6)
    np_ = xyz.shape[0]
    hinges = _hinges()
    folded = 0
    for s, k, dist in PUSH:
        x = pos[s * np_:(s + 1) * np_]
        rel = x - x[0]
        rel -= box * np.round(rel / box)  # unwrapped around bead 0
        d = rel.mean(axis=0) - rel[k]
        x[k] += dist * d / np.linalg.norm(d)
        rel[k] += dist * d / np.linalg.norm(d)
        folded += int((_fold_cosines(rel, hinges) < 0).sum())
    return pos, vel, folded


def _system(brownian=False):
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=20, brownian=1 if brownian else 0, **CFG0)
    p.skin = 0.3
    return p, sg.System(t, 20, p)


def test_folded_hinge_forces_and_energies():
    p, s = _system()
    pos, vel, folded = _folded_state(p.box, 51)
    assert folded >= 3
    o = Oracle(20, **CFG0)
    o.set_state(pos, vel)
    ref = o.forces()
    s.set_state(pos, vel)
    got = s.forces()
    for name, g, r in zip(["total", "stretch", "bend", "yukawa", "lj"], got, ref):
        if np.linalg.norm(r) > 0:
            assert rel_norm(g, r) < 1e-10, (name, rel_norm(g, r))
    np.testing.assert_allclose(s.energies()[0], o.energies(), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("brownian", [False, True])
def test_folded_hinge_trajectory(brownian):
    p, s = _system(brownian)
    pos, vel, folded = _folded_state(p.box, 52)
    assert folded >= 3
    o = Oracle(20, brownian=brownian, L=3.0, B=5, **CFG0)
    o.set_state(pos, vel)
    s.set_state(pos, vel)
    o.run(100)
    s.step(100)
    op, ov, _, _ = o.get_state()
    gp, gv = s.get_state()
    assert rel_norm(gp, op) < 1e-8 and rel_norm(gv, ov) < 1e-8, (rel_norm(gp, op), rel_norm(gv, ov))
