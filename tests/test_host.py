"""Host-side logic of the GPU library (no device calls): exported C ABI, template reader and
derived topology, parameter derivation, initial-state generator."""
import os
import re

import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import REPO, TEMPLATE, TEMPLATE_BIN, Oracle


def _header_functions():
    text = open(os.path.join(REPO, "include", "souffle_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_\w+)\s*\(", text)))


def test_library_exports_every_header_function():
    lib = sg.load()
    names = _header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in sg.SIGNATURES, n


@pytest.mark.parametrize("path", [TEMPLATE, TEMPLATE_BIN])
def test_template_topology(path):
    t = sg.Template(path)
    assert (t.n_beads, t.n_edges, t.n_faces, t.n_hinges) == (105, 333, 260, 156)
    assert t.centre == 42
    assert t.n_charged == (6 if path == TEMPLATE else 0)
    # the oracle associates faces to edges by the reference's own scan; both must agree
    o = Oracle(1, template=path, C_uM=10.0)
    assert np.array_equal(t.hinges(), o.hinges())


def test_template_rejects_missing_file(tmp_path):
    with pytest.raises(sg.SfError):
        sg.Template(str(tmp_path / "nope"))


# Appendix-D style values computed from the reference's formulas (src/md.cpp:159-173)
@pytest.mark.parametrize("path,ns,c,box,kappa,ecut,ecut_el", [
    (TEMPLATE, 20, 100.0, 34.626, 1.03003, 2.5, 11.650),
    (TEMPLATE, 2000, 100.0, 160.719, 1.03003, 2.5, 11.650),
    (TEMPLATE, 5000, 100.0, 218.129, 1.03003, 2.5, 11.650),
    (TEMPLATE, 20, 800.0, 34.626, 2.91336, 2.5, 4.119),
    (TEMPLATE_BIN, 20, 100.0, 38.954, 0.91559, 2.2222, 13.106),
    (TEMPLATE_BIN, 2000, 100.0, 180.806, 0.91559, 2.2222, 13.106),
])
def test_params_derive(path, ns, c, box, kappa, ecut, ecut_el):
    p = sg.Template(path).params(C_uM=800.0, c_mM=c, n_subunits=ns)
    assert abs(p.box - box) < 1e-3
    assert abs(p.kappa - kappa) < 1e-5
    assert abs(p.ecut - ecut) < 1e-4
    assert abs(p.ecut_el - ecut_el) < 1e-3
    o = Oracle(ns if ns <= 20 else 20, template=path, C_uM=800.0, c_mM=c)
    assert p.kappa == o.kappa and p.ecut_el == o.ecut_el and p.lb == o.lb and p.ecut == o.ecut


def test_make_state():
    t = sg.Template(TEMPLATE)
    p = t.params(C_uM=800.0, c_mM=100.0, n_subunits=200)
    pos, vel = t.make_state(200, p.box, 1.3, 17)
    assert np.all(np.abs(pos) <= p.box / 2 + 1e-12)
    ke = 0.5 * np.sum(vel * vel)
    assert abs(ke / (1.5 * len(vel)) - 1.3) < 1e-12
    assert np.allclose(vel.sum(axis=0), 0, atol=1e-9)
    # subunits are rigid template copies and beads of different subunits are >= 1 sigma apart
    xyz = np.array([[float(v) for v in l.split()[1:4]] for l in open(TEMPLATE).read().split("mass\n")[1].split("\n")[:105]])
    d0 = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    sub = pos[:105]
    dd = sub[:, None] - sub[None]
    dd = np.where(dd > p.box / 2, dd - p.box, np.where(dd < -p.box / 2, dd + p.box, dd))
    assert np.allclose(np.linalg.norm(dd, axis=-1), d0, atol=1e-9)
    a, b = pos[:105 * 20], pos[105 * 20:105 * 40]
    d = a[:, None] - b[None]
    d = np.where(d > p.box / 2, d - p.box, np.where(d < -p.box / 2, d + p.box, d))
    assert np.sqrt((d * d).sum(-1)).min() >= 1.0 - 1e-12
    pos2, vel2 = t.make_state(200, p.box, 1.3, 17)
    assert np.array_equal(pos, pos2) and np.array_equal(vel, vel2)


def test_big_bead_placement_matches_oracle():
    """sf_big_beads_place against the oracle's generate_big_beads restatement
    (src/initialize.cpp:566-680): the same rand() stream and long double arithmetic, so the
    positions agree bit for bit; too few attempts is an error, as the reference's exit(1)."""
    from oracle_lib import place_big_beads as orc_place
    from souffle_states import make_state, read_template_xyz
    t = sg.Template(TEMPLATE)
    o = Oracle(20)
    pos, _ = make_state(read_template_xyz(TEMPLATE)[0], 20, o.box, 3, spacing=8.6, perturb=0.05)
    sig, _ = o.bead_props()
    for n in (1, 4):
        got = sg.place_big_beads(t, pos, o.box, n)
        ref = orc_place(pos, sig, o.box, n, 6.0)
        assert np.array_equal(got, ref)
    with pytest.raises(sg.SfError):
        sg.place_big_beads(t, pos, o.box, 3, max_attempts=1)
