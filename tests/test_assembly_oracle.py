"""Pins the oracle's oligomer (mass-spectrum) analysis, oracle/souffle_oracle.cpp `oligomers`
(restating src/analysis.cpp:143-186), CPU only:
  - against connected components (scipy) of the brute-force bead-contact graph of the same
    positions, for random and MD-evolved states, both precisions;
  - on constructed rows of subunits with known contacts (labels and mass spectrum by hand),
    including a contact across the periodic boundary and a chain whose breadth-first growth
    has to come back to lower-numbered subunits.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from oracle_lib import TEMPLATE, Oracle
from souffle_states import make_state, min_bead_distance, read_template_xyz, subunit_line

XYZ = read_template_xyz(TEMPLATE)[0]


def brute_oligomers(pos, ns, box, contact=2.0):
    np_ = len(pos) // ns
    adj = np.zeros((ns, ns), dtype=bool)
    for s in range(ns):
        for t in range(s + 1, ns):
            adj[s, t] = min_bead_distance(pos[s * np_:(s + 1) * np_], pos[t * np_:(t + 1) * np_], box) < contact
    _, comp = connected_components(csr_matrix(adj), directed=False)
    label = np.array([np.nonzero(comp == comp[s])[0].min() for s in range(ns)])
    ms = np.zeros(ns, dtype=int)
    for c in np.unique(comp):
        ms[(comp == c).sum() - 1] += 1
    return label, ms


def test_oligomers_match_connected_components():
    o = Oracle(20, L=0.3, B=5, elj_att=4.0)
    old = Oracle(20, L=0.3, B=5, elj_att=4.0, long_double=True)
    pos, vel = make_state(XYZ, 20, o.box, 5, T=o.T)
    o.set_state(pos, vel)
    seen = set()
    for k in range(4):
        if k:
            o.run(800)
        p = o.get_state()[0]
        lab, ms, n = o.oligomers(2.0)
        blab, bms = brute_oligomers(p, 20, o.box)
        assert np.array_equal(lab, blab) and np.array_equal(ms, bms)
        assert n == bms.sum() and (np.arange(1, 21) * ms).sum() == 20
        seen.add(n)
        # the long double restatement on the same positions
        old.set_state(p)
        assert np.array_equal(old.oligomers(2.0)[0], lab)
        # a larger contact distance joins at least as much
        lab3, ms3, n3 = o.oligomers(3.0)
        assert np.array_equal(lab3, brute_oligomers(p, 20, o.box, 3.0)[0]) and n3 <= n
    assert len(seen) > 1  # oligomers formed during the run


def test_oligomers_constructed_rows():
    o = Oracle(8, C_uM=5.0, L=0.3, B=5)
    box = o.box
    assert box > 120
    # 0-1-2 touch, 3-4 touch, 5 alone; then 6-7 touching across the +x / -x boundary
    row = subunit_line(XYZ, [True, True, False, True, False], box, 3, origin=[-45.0, 0, 0])
    pair = subunit_line(XYZ, [True], box, 4, origin=[box / 2 - 6.0, 20.0, 20.0])
    pos = np.concatenate([row, pair])
    assert (np.abs(pos) <= box / 2).all() and np.ptp(pair[:, 0]) > box / 2  # the pair is split by the wrap
    o.set_state(pos)
    lab, ms, n = o.oligomers(2.0)
    assert lab.tolist() == [0, 0, 0, 3, 3, 5, 6, 6]
    assert ms.tolist() == [1, 2, 1, 0, 0, 0, 0, 0] and n == 4
    assert np.array_equal(lab, brute_oligomers(pos, 8, box)[0])


def test_oligomers_growth_returns_to_lower_subunits():
    """Subunit order 0, 2, 1 along the row: 0 touches 2 only, 1 touches 2 only. The oligomer seeded
    at 0 must reach 1 through 2 (src/analysis.cpp:154-177 grows over the members added in each
    round and scans all higher-numbered subunits)."""
    o = Oracle(3, C_uM=20.0, L=0.3, B=5)
    row = subunit_line(XYZ, [True, True], o.box, 9)
    np_ = len(XYZ)
    pos = np.concatenate([row[:np_], row[2 * np_:], row[np_:2 * np_]])  # subunit 1 is the far end
    o.set_state(pos)
    lab, ms, n = o.oligomers(2.0)
    assert lab.tolist() == [0, 0, 0] and ms.tolist() == [0, 0, 1] and n == 1
    lab, ms, n = o.oligomers(1.2)  # below the constructed contact distance: three monomers
    assert lab.tolist() == [0, 1, 2] and n == 3
