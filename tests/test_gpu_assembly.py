"""Oligomer (mass-spectrum) analysis on the GPU (sf_oligomers) against the oracle's restatement
of src/analysis.cpp:143-186, and long-run statistics of the GPU path against the oracle's.

- labels and mass spectra are integers: bit-exact equality with the oracle on the same
  positions (constructed rows with known answers, MD-evolved states, a replica batch), and with
  a brute-force contact graph at 2000 subunits;
- long-run statistics: an ensemble of 16 seeded 20-subunit systems (the small-system
  parameters, 10,000 steps) on the GPU against the oracle's ensemble from the same initial
  states (tests/golden/assembly_stats.json, written by tests/golden/make_assembly_stats.py).
  Trajectories diverge chaotically, so ensemble means are compared: temperature and potential
  energy per bead over steps >= 2000, number of oligomers and largest oligomer over steps
  >= 5000, and for MD the extended-energy change. Each must agree within 4 combined standard
  errors of the seed means (GPU FP64 MD, GPU FP32-pair MD, GPU BD).
"""
import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

import souffle_gpu as sg
from oracle_lib import REPO, TEMPLATE, Oracle
from souffle_states import make_state, min_bead_distance, read_template_xyz, subunit_line

pytestmark = pytest.mark.gpu

XYZ = read_template_xyz(TEMPLATE)[0]
FIXTURE = os.path.join(REPO, "tests", "golden", "assembly_stats.json")


@pytest.fixture(scope="module")
def template():
    return sg.Template(TEMPLATE)


def _system(template, ns, pos, vel=None, **kw):
    cli = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
    cli.update(kw)
    p = template.params(n_subunits=ns, **cli)
    s = sg.System(template, ns, p)
    s.set_state(pos, vel)
    return s, p


def test_oligomers_constructed_rows(template):
    o = Oracle(8, C_uM=5.0, L=0.3, B=5)
    box = o.box
    row = subunit_line(XYZ, [True, True, False, True, False], box, 3, origin=[-45.0, 0, 0])
    pair = subunit_line(XYZ, [True], box, 4, origin=[box / 2 - 6.0, 20.0, 20.0])
    pos = np.concatenate([row, pair])
    s, p = _system(template, 8, pos, C_uM=5.0)
    assert abs(p.box - box) < 1e-9 * box
    (lab, ms), = s.oligomers(2.0)
    assert lab.tolist() == [0, 0, 0, 3, 3, 5, 6, 6]
    assert ms.tolist() == [1, 2, 1, 0, 0, 0, 0, 0]
    # subunit 1 at the far end of a row 0 - 2 - 1
    np_ = len(XYZ)
    o3 = Oracle(3, C_uM=20.0, L=0.3, B=5)
    row = subunit_line(XYZ, [True, True], o3.box, 9)
    pos3 = np.concatenate([row[:np_], row[2 * np_:], row[np_:2 * np_]])
    s3, _ = _system(template, 3, pos3, C_uM=20.0)
    assert s3.oligomers(2.0)[0][0].tolist() == [0, 0, 0]
    assert s3.oligomers(1.2)[0][0].tolist() == [0, 1, 2]


def test_oligomers_match_oracle_along_md(template):
    o = Oracle(20, L=0.3, B=5, elj_att=4.0)
    pos, vel = make_state(XYZ, 20, o.box, 5, T=o.T)
    s, _ = _system(template, 20, pos, vel, elj_att=4.0)
    counts = set()
    for k in range(5):
        if k:
            s.step(800)
        gp = s.get_state()[0]
        o.set_state(gp)
        for contact in (2.0, 3.0):
            (lab, ms), = s.oligomers(contact)
            olab, oms, n = o.oligomers(contact)
            assert np.array_equal(lab, olab) and np.array_equal(ms, oms), (k, contact)
        counts.add(n)
    assert len(counts) > 1


def test_oligomers_replica_batch(template):
    sizes, seeds = [20, 7, 33], [11, 12, 13]
    ps, poss, vels = [], [], []
    for n, sd in zip(sizes, seeds):
        p = template.params(n_subunits=n, C_uM=800.0, c_mM=100.0, elj_att=4.0, dt=0.005, K=298.0)
        pos, vel = make_state(XYZ, n, p.box, sd, T=p.T)
        ps.append(p); poss.append(pos); vels.append(vel)
    s = sg.System(template, sizes, ps)
    s.set_state(np.concatenate(poss), np.concatenate(vels))
    s.step(1500)
    gp = s.get_state()[0]
    out = s.oligomers(2.0)
    b0 = 0
    for r, n in enumerate(sizes):
        o = Oracle(n, L=0.3, B=5, elj_att=4.0)
        o.set_state(gp[b0:b0 + n * len(XYZ)])
        olab, oms, _ = o.oligomers(2.0)
        assert np.array_equal(out[r][0], olab) and np.array_equal(out[r][1], oms), r
        b0 += n * len(XYZ)


def test_oligomers_2000_subunits_brute_force(template):
    ns = 2000
    p = template.params(n_subunits=ns, C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
    pos, vel = template.make_state(ns, p.box, p.T, 2000)
    s = sg.System(template, ns, p)
    s.set_state(pos, vel)
    s.step(300)
    gp = s.get_state()[0]
    np_ = len(XYZ)
    sub = gp.reshape(ns, np_, 3)
    cen = sub[:, template.centre, :]
    box = p.box
    ext = 0.0
    for k in range(ns):
        d = sub[k] - cen[k]
        d = np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))
        ext = max(ext, np.sqrt((d * d).sum(axis=1).max()))
    reach = 3.0 + 2 * ext + 1e-6
    pairs = []  # (a, b, closest bead distance) of subunits whose centres are within reach
    for a in range(ns):
        d = cen[a + 1:] - cen[a]
        d = np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))
        for b in np.nonzero((d * d).sum(axis=1) < reach * reach)[0] + a + 1:
            pairs.append((a, b, min_bead_distance(sub[a], sub[b], box)))
    for contact in (2.0, 3.0):
        (lab, ms), = s.oligomers(contact)
        e = [(a, b) for a, b, dm in pairs if dm < contact]
        adj = csr_matrix((np.ones(len(e)), ([x[0] for x in e], [x[1] for x in e])), shape=(ns, ns))
        _, comp = connected_components(adj, directed=False)
        first = {}
        for k in range(ns):
            first.setdefault(comp[k], k)
        assert lab.tolist() == [first[c] for c in comp], contact
        sizes = np.bincount(comp)
        assert ms.tolist() == np.bincount(sizes, minlength=ns + 1)[1:].tolist(), contact
        assert ms.sum() < ns  # contacts exist at this density


# ---------------------------------------------------------------- long-run statistics

def _stats(rows, n_beads):
    rows = np.asarray(rows)
    step = rows[:, 0]
    late, later = step >= 2000, step >= 5000
    T = rows[late, 1].mean() / (1.5 * n_beads)
    pe = rows[late, 2:6].sum(axis=1).mean() / n_beads
    h = rows[:, 1:8].sum(axis=1)
    return {"T": T, "pe": pe, "n_olig": rows[later, 8].mean(), "largest": rows[later, 9].mean(),
            "dH": (h[-1] - h[0]) / n_beads}


def _ensemble(runs, n_beads):
    per = [_stats(r, n_beads) for r in runs]
    out = {}
    for k in per[0]:
        v = np.array([x[k] for x in per])
        out[k] = (v.mean(), v.std(ddof=1) / np.sqrt(len(v)))
    return out


def _gpu_runs(template, fx, brownian, fp32):
    ns, seeds, P = fx["n_subunits"], fx["seeds"], fx["params"]
    p = template.params(n_subunits=ns, C_uM=P["C_uM"], c_mM=P["c_mM"], elj_att=P["elj_att"], dt=P["dt"],
                        K=P["K"], ks=P["ks"], kb=P["kb"], ecut_c=P["ecut_c"], Q=P["Q"], q_chain=P["q_chain"],
                        damp=P["damp"], brownian=1 if brownian else 0)
    p.pair_fp32 = 1 if fp32 else 0
    poss, vels = zip(*[make_state(XYZ, ns, p.box, sd, T=p.T) for sd in seeds])
    s = sg.System(template, [ns] * len(seeds), [p] * len(seeds))
    s.set_state(np.concatenate(poss), np.concatenate(vels))
    runs = [[] for _ in seeds]
    for k in range(fx["steps"] // fx["every"] + 1):
        if k:
            s.step(fx["every"])
        e = s.energies()
        olig = s.oligomers(2.0)
        for r in range(len(seeds)):
            lab, ms = olig[r]
            runs[r].append([k * fx["every"]] + e[r].tolist() + [int(ms.sum()), int(np.nonzero(ms)[0].max()) + 1])
    return runs


@pytest.mark.parametrize("form", ["md_fp64", "md_fp32_pairs", "bd_fp64"])
def test_long_run_statistics_match_oracle_ensemble(template, form):
    fx = json.load(open(FIXTURE))
    brownian = form.startswith("bd")
    cpu = fx["bd" if brownian else "md"]
    nb = fx["n_subunits"] * len(XYZ)
    # the fixture's first rows are the shared initial states
    gpu = _gpu_runs(template, fx, brownian, form.endswith("fp32_pairs"))
    for g, c in zip(gpu, cpu):
        assert np.allclose(g[0][1:8], c[0][1:8], rtol=1e-10, atol=1e-9) and g[0][8:] == c[0][8:]
    G, C = _ensemble(gpu, nb), _ensemble(cpu, nb)
    keys = ["T", "pe", "n_olig", "largest"] + ([] if brownian else ["dH"])
    for k in keys:
        (mg, sg_), (mc, sc) = G[k], C[k]
        print(f"{form} {k}: gpu {mg:.5g} +- {sg_:.2g}, oracle {mc:.5g} +- {sc:.2g}")
        assert abs(mg - mc) <= 4 * np.hypot(sg_, sc) + 1e-12, (k, G[k], C[k])
    # oligomers formed in both ensembles
    assert C["n_olig"][0] < fx["n_subunits"] - 1 and G["n_olig"][0] < fx["n_subunits"] - 1
