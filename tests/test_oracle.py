"""Pins the CPU oracle (oracle/souffle_oracle.cpp) without the reference binary.

The reference's hot-path sources cannot be compiled here (they need Boost), and the repository
ships no tests or golden vectors for this path, so the oracle is pinned by what the mathematics
fixes, each check chosen so that a plausible transcription mistake fails it:
  - pair forces against -dU/dr of the closed-form potentials (central differences), including
    the asymmetric rules: sigma_i in the attractive LJ force, epsilon = 50 for type_j == 7,
    ordered lj_a rows, Yukawa size factor and cutoff;
  - energies against the closed forms with their cutoff shifts;
  - bending against E = Kb (1 - |n0 . n1|) computed from numpy face normals (value and
    finite-difference gradient), for both face orientations; past a 90 degree fold, where the
    reference's force is no longer that gradient, against its closed form built from
    finite-difference gradients of the face areas and the unflipped area product;
  - stretching against E = ks/2 (len0 - L)^2;
  - the whole trimer_MVM system: force = -grad(se + be + ne + ce) by finite differences;
  - Nose-Hoover: energy conservation with q = 1 and conservation of the extended energy with
    q = 5; BD: fluctuation-dissipation (free beads heat to kT = 1 at the drag rate);
  - the BD noise stream against an independent Python minstd_rand0 + generate_canonical;
  - the reference neighbour list contains every pair inside the force cutoffs;
  - the FP64 mode agrees with the x87 long double mode to round-off;
  - a run continued in a fresh oracle from (pos, vel, xi, eta, particle KE) follows the
    uninterrupted run to round-off (orc_set_thermostat, used to continue GPU states).
"""
import math
import os

import numpy as np
import pytest

from oracle_lib import TEMPLATE, TEMPLATE_BIN, Oracle, bd_uniforms, rel_norm
from souffle_states import make_state, read_template_xyz
from template_writer import write_template

KAPPA100 = math.sqrt(8 * 3.14159 * (100 * 6.022e23 * 1e-27) * 0.701)


def _deriv(f, r, h=1e-4):
    """Fourth-order central difference of a scalar function."""
    return (-f(r + 2 * h) + 8 * f(r + h) - 8 * f(r - h) + f(r - 2 * h)) / (12 * h)


def _pair_system(tmp_path, bead_i, bead_j, attractions, r, c_mM=100.0, elj_att=2.1, name="pair"):
    """Two subunits of a two-bead template (template bead 0 = j's parameters, bead 1 = i's).
    Bead i = subunit 0's bead 1 (global 1) sits at distance r from bead j = subunit 1's bead 0
    (global 2); every other pair is > 19 sigma apart. Subunit 0's centre bead is its bead 0,
    so the probed pair avoids the reference's stale-r2 rule for the centre bead
    (src/functions.cpp:97)."""
    path = write_template(str(tmp_path / name), beads=[bead_j, bead_i], attractions=attractions)
    o = Oracle(2, template=path, C_uM=10.0, c_mM=c_mM, elj_att=elj_att, L=7.5, B=1)
    d = np.array([0.3, -0.5, 0.81])
    d = d / np.linalg.norm(d) * r
    pos = np.array([[0.0, 0.0, 20.0], [0.0, 0.0, 0.0], d, [d[0], d[1], d[2] - 20.0]])
    o.set_state(pos)
    assert o.centre_local() == 0
    f = [a[[1, 0, 3, 2]] for a in o.forces()]  # rows reordered: 0 = i, 3 = j
    return o, f, -d  # r_vec = x_i - x_j


def _lj_force(r_vec, eps, s):
    r = np.linalg.norm(r_vec)
    U = lambda x: 4 * eps * ((s / x) ** 12 - (s / x) ** 6)
    return -_deriv(U, r) * r_vec / r


@pytest.mark.parametrize("r", [0.95, 1.2, 1.9, 2.4])
def test_attractive_lj_uses_sigma_i(tmp_path, r):
    # types (2, 2) attractive; sigma_i = 1.0, sigma_j = 1.3 (src/forces.cpp:126-131)
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 2, 0.0, 1.0, 1.0), (0, 0, 0, 2, 0.0, 1.3, 1.0), [(2, 2)], r)
    F = f[4]
    np.testing.assert_allclose(F[0], _lj_force(rv, 2.1, 1.0), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(F[3], _lj_force(-rv, 2.1, 1.3), rtol=1e-9, atol=1e-12)
    assert np.all(F[1] == 0) and np.all(F[2] == 0)


def test_attractive_lj_cutoff(tmp_path):
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 2, 0.0, 1.0, 1.0), (0, 0, 0, 2, 0.0, 1.0, 1.0), [(2, 2)], 2.51)
    assert np.all(f[4] == 0)


def test_type7_epsilon_50(tmp_path):
    # rows (2,7) and (7,2): the force on the type-2 bead (j of type 7) uses eps = 50
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 2, 0.0, 1.0, 1.0), (0, 0, 0, 7, 0.0, 1.0, 1.0), [(2, 7), (7, 2)], 1.3)
    np.testing.assert_allclose(f[4][0], _lj_force(rv, 50.0, 1.0), rtol=1e-9)
    np.testing.assert_allclose(f[4][3], _lj_force(-rv, 2.1, 1.0), rtol=1e-9)


def test_lj_rows_are_ordered(tmp_path):
    # only (1, 2) attractive: bead i (type 1, j type 2) attracts; bead j (type 2, i type 1) is WCA
    r = 1.05
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 1, 0.0, 1.0, 1.0), (0, 0, 0, 2, 0.0, 1.0, 1.0), [(1, 2)], r)
    np.testing.assert_allclose(f[4][0], _lj_force(rv, 2.1, 1.0), rtol=1e-9)
    wca = _lj_force(-rv, 1.0, 1.0)
    np.testing.assert_allclose(f[4][3], wca, rtol=1e-9)
    assert not np.allclose(f[4][0], -f[4][3])


@pytest.mark.parametrize("r,on", [(0.9, True), (1.1, True), (1.2, False)])
def test_wca(tmp_path, r, on):
    # no attractive row: WCA with sigma = (s_i + s_j)/2 = 1.0, eps = 1, cut at 1.12246205 sigma
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 1, 0.0, 0.9, 1.0), (0, 0, 0, 1, 0.0, 1.1, 1.0), [(2, 2)], r)
    if on:
        np.testing.assert_allclose(f[4][0], _lj_force(rv, 1.0, 1.0), rtol=1e-9)
        np.testing.assert_allclose(f[4][3], _lj_force(-rv, 1.0, 1.0), rtol=1e-9)
    else:
        assert np.all(f[4] == 0)


@pytest.mark.parametrize("r", [1.0, 3.0, 8.0, 11.0, 12.0])
def test_yukawa(tmp_path, r):
    si, sj, qi, qj = 1.0, 1.3, 1.0, -2.0
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 1, qi, si, 1.0), (0, 0, 0, 1, qj, sj, 1.0), [], r)
    k, lb = o.kappa, o.lb
    assert abs(k - KAPPA100) < 1e-12
    shc = (si + sj) / 2
    Y = math.exp(k * shc) / ((1 + k * si / 2) * (1 + k * sj / 2))
    U = lambda x: Y * qi * qj * lb * math.exp(-k * x) / x
    expect = -_deriv(U, r) * rv / r
    if r < o.ecut_el:
        np.testing.assert_allclose(f[3][0], expect, rtol=1e-9)
        np.testing.assert_allclose(f[3][3], -expect, rtol=1e-9)
    else:
        assert np.all(f[3][0] == 0)
    # pair energy (attributed to the lower id, shifted at ecut_el; src/energies.cpp:54-56)
    tot, pb = o.energies(per_bead=True)
    eu = (U(r) - U(o.ecut_el)) if r < o.ecut_el else 0.0
    assert abs(tot[4] - eu) <= 1e-12 * max(1.0, abs(eu))
    assert abs(pb[1, 3] - eu) <= 1e-12 * max(1.0, abs(eu))  # bead i = global 1 < j = global 2


def _yukawa_all_pairs(pos, q, sig, box, kappa, lb):
    """Closed-form uncut Yukawa over every charged pair i != j, single-shift minimum image:
    forces [N,3] and the total (unshifted) energy."""
    n = len(pos)
    f = np.zeros((n, 3))
    e = 0.0
    for i in range(n):
        for j in range(n):
            if i == j or q[i] == 0 or q[j] == 0:
                continue
            r = pos[i] - pos[j]
            r = np.where(r > box / 2, r - box, np.where(r < -box / 2, r + box, r))
            d = np.linalg.norm(r)
            shc = (sig[i] + sig[j]) / 2
            Y = math.exp(kappa * shc) / ((1 + kappa * sig[i] / 2) * (1 + kappa * sig[j] / 2))
            U = lambda x: Y * q[i] * q[j] * lb * math.exp(-kappa * x) / x
            f[i] += Y * q[i] * q[j] * lb * math.exp(-kappa * d) / d ** 2 * (kappa + 1 / d) * r
            if i < j:
                e += U(d)
    return f, e


@pytest.mark.parametrize("c_mM", [10.0, 100.0])
def test_uncut_yukawa_all_pairs(tmp_path, c_mM):
    """-e 0 with a list radius covering the box: every charged pair interacts (minimum image,
    intra-subunit pairs included), energies unshifted (src/forces.cpp:104-107,
    src/energies.cpp:54-58)."""
    beads = [(0, 0, 0, 1, 0.0, 1.0, 1.0), (1.1, 0, 0, 1, 1.0, 1.0, 1.0), (0, 1.2, 0.3, 1, -2.0, 1.3, 1.0),
             (-0.9, -0.4, 0.5, 1, 1.0, 1.1, 1.0)]
    path = write_template(str(tmp_path / "q3"), beads=beads)
    ns = 6
    o = Oracle(ns, template=path, C_uM=50.0, c_mM=c_mM, ecut_c=0.0, L=80.0, B=1)
    assert o.ecut_el == 0.0
    rng = np.random.default_rng(5)
    body = np.array([b[:3] for b in beads], dtype=float)
    cen = []
    while len(cen) < ns:
        c = rng.uniform(-o.box / 2 + 2, o.box / 2 - 2, 3)
        if all(np.linalg.norm(c - x) > 6.0 for x in cen):
            cen.append(c)
    pos = np.concatenate([body + c for c in cen])
    o.set_state(pos)
    fy = o.forces()[3]
    q = np.tile([b[4] for b in beads], ns)
    sig = np.tile([b[5] for b in beads], ns)
    f_ref, e_ref = _yukawa_all_pairs(pos, q, sig, o.box, o.kappa, o.lb)
    assert rel_norm(fy, f_ref) < 1e-12
    assert abs(o.energies()[4] - e_ref) <= 1e-12 * abs(e_ref)
    # the default list radius (ecut + 7.5) leaves the far pairs out, as the reference does
    o2 = Oracle(ns, template=path, C_uM=50.0, c_mM=c_mM, ecut_c=0.0, L=7.5, B=1)
    o2.set_state(pos)
    assert rel_norm(o2.forces()[3], f_ref) > 1e-9


@pytest.mark.parametrize("r", [0.95, 1.5, 2.4])
def test_lj_energy_shifted(tmp_path, r):
    # attractive energy uses s = (s_i + s_j)/2 and the shift at ecut (src/energies.cpp:63-86)
    o, f, rv = _pair_system(tmp_path, (0, 0, 0, 2, 0.0, 1.0, 1.0), (0, 0, 0, 2, 0.0, 1.2, 1.0), [(2, 2)], r)
    s, e, ec = 1.1, 2.1, o.ecut
    U = lambda x: 4 * e * ((s / x) ** 12 - (s / x) ** 6)
    np.testing.assert_allclose(o.energies()[3], U(r) - U(ec), rtol=1e-12)


def _hinge_energy(x, faces, kb):
    def normal(a, b, c):
        A = np.cross(x[b] - x[a], x[c] - x[b])
        return A / np.linalg.norm(A)
    n0, n1 = normal(*faces[0]), normal(*faces[1])
    return kb * (1 - abs(np.dot(n0, n1)))


@pytest.mark.parametrize("faces", [[(0, 1, 2), (0, 1, 3)], [(0, 1, 2), (1, 0, 3)], [(2, 1, 0), (0, 3, 1)]])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bending_hinge_closed_form(tmp_path, faces, seed):
    rng = np.random.default_rng(seed)
    base = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.9, 0.0], [0.5, -0.9, 0.0]])
    beads = [(*b, 1, 0.0, 1.0, 1.0) for b in base]
    edges = [(0, 1, 1, 1.0), (0, 2, 0, 1.0), (1, 2, 0, 1.0), (0, 3, 0, 1.0), (1, 3, 0, 1.0)]
    path = write_template(str(tmp_path / "hinge"), beads, edges, faces, [])
    kb = 20.0
    o = Oracle(1, template=path, C_uM=10.0, ks=0.0, kb=kb, L=7.5)
    x = base + rng.normal(scale=0.25, size=base.shape)
    x[3, 2] += rng.choice([-0.6, 0.6])  # fold out of plane either way
    o.set_state(x)
    tot = o.energies()
    assert abs(tot[2] - _hinge_energy(x, faces, kb)) < 1e-12 * kb
    F = o.forces()[2]
    h = 1e-6
    for i in range(4):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            g = (_hinge_energy(xp, faces, kb) - _hinge_energy(xm, faces, kb)) / (2 * h)
            assert abs(F[i, c] + g) < 1e-7 * kb, (i, c, F[i, c], -g)
    assert np.allclose(F.sum(axis=0), 0, atol=1e-12)


@pytest.mark.parametrize("faces", [[(0, 1, 2), (0, 1, 3)], [(0, 1, 2), (1, 0, 3)], [(2, 1, 0), (0, 3, 1)]])
@pytest.mark.parametrize("angle", [60.0, 100.0, 150.0, 210.0, 300.0])
def test_bending_folded_hinge(tmp_path, faces, angle):
    """The reference's hinge force at any fold (src/edge.cpp:69-188): Kb/(4 a0 a1) (grad P -
    |S| (grad a0/a0 + grad a1/a1)) with P = -(e x a).(e x b) the product its gradient formulas
    differentiate (e = x1-x0, a = x2-x0, b = x3-x0; independent of the faces' winding), S = 4 A0.A1
    and a0, a1 the face areas; gradients here by central differences of those scalars. Up to a 90
    degree fold this is minus the gradient of Kb (1 - |n0.n1|); past it, it is not."""
    base = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.9, 0.0], [0.5, -0.9, 0.0]])
    beads = [(*b, 1, 0.0, 1.0, 1.0) for b in base]
    edges = [(0, 1, 1, 1.0), (0, 2, 0, 1.0), (1, 2, 0, 1.0), (0, 3, 0, 1.0), (1, 3, 0, 1.0)]
    path = write_template(str(tmp_path / "hinge"), beads, edges, faces, [])
    kb = 20.0
    o = Oracle(1, template=path, C_uM=10.0, ks=0.0, kb=kb, L=7.5)
    t = math.radians(angle)
    x = base.copy()
    x[3] = [0.5, -0.9 * math.cos(t), -0.9 * math.sin(t)]  # bead 3 turned about the hinge axis
    x = x + np.random.default_rng(int(angle)).normal(scale=0.03, size=x.shape)
    o.set_state(x)
    F = o.forces()[2]

    def uv(y):
        e, a, b = y[1] - y[0], y[2] - y[0], y[3] - y[0]
        return np.cross(e, a), np.cross(e, b)

    def grad(f, y, h=1e-6):
        g = np.zeros_like(y)
        for i in range(4):
            for c in range(3):
                yp, ym = y.copy(), y.copy()
                yp[i, c] += h
                ym[i, c] -= h
                g[i, c] = (f(yp) - f(ym)) / (2 * h)
        return g

    u, v = uv(x)
    nu, nv = np.linalg.norm(u), np.linalg.norm(v)
    gP = grad(lambda y: -np.dot(*uv(y)), x)
    ga0 = grad(lambda y: np.linalg.norm(uv(y)[0]), x)
    ga1 = grad(lambda y: np.linalg.norm(uv(y)[1]), x)
    expect = kb / (nu * nv) * (gP - abs(np.dot(u, v)) * (ga0 / nu + ga1 / nv))
    np.testing.assert_allclose(F, expect, rtol=0, atol=2e-7 * kb)
    # the energy gradient agrees below a 90 degree fold only
    c = -np.dot(u, v) / (nu * nv)
    gE = grad(lambda y: kb * (1 - abs(np.dot(*uv(y))) / (np.linalg.norm(uv(y)[0]) * np.linalg.norm(uv(y)[1]))), x)
    assert (np.abs(F + gE).max() < 1e-6 * kb) == (c > 0)


def test_stretching_closed_form(tmp_path):
    beads = [(0, 0, 0, 1, 0.0, 1.0, 1.0), (1, 0, 0, 1, 0.0, 1.0, 1.0)]
    path = write_template(str(tmp_path / "bond"), beads, [(0, 1, 0, 1.0)], [], [])
    ks = 50.0
    o = Oracle(1, template=path, C_uM=10.0, ks=ks, L=7.5)
    x = np.array([[0.1, 0.2, -0.1], [1.2, 0.1, 0.3]])
    o.set_state(x)
    L = np.linalg.norm(x[0] - x[1])
    np.testing.assert_allclose(o.forces()[1][0], ks * (1.0 - L) / L * (x[0] - x[1]), rtol=1e-13)
    assert abs(o.energies()[1] - 0.5 * ks * (1.0 - L) ** 2) < 1e-13


@pytest.mark.parametrize("long_double", [False, True])
def test_trimer_force_is_minus_energy_gradient(long_double):
    xyz = read_template_xyz(TEMPLATE)[0]
    o = Oracle(20, long_double=long_double)
    pos, vel = make_state(xyz, 20, o.box, 3, spacing=8.6, perturb=0.05)
    o.set_state(pos, vel)
    F = o.forces()[0]

    def epot(p):
        o.set_state(p, vel)
        e = o.energies()
        return e[1] + e[2] + e[3] + e[4]
    rng = np.random.default_rng(5)
    # beads with LJ contacts, charged beads and random beads
    flj = np.linalg.norm(o.forces()[4], axis=1)
    picks = list(np.argsort(flj)[-4:]) + [99, 104, 105 * 7 + 101] + list(rng.choice(len(pos), 4, replace=False))
    o.set_state(pos, vel)
    h = 1e-5
    for i in picks:
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            g = (epot(pp) - epot(pm)) / (2 * h)
            assert abs(F[i, c] + g) < 2e-7 * max(1.0, abs(F[i, c])), (i, c, F[i, c], -g)


def test_fp64_matches_long_double():
    xyz = read_template_xyz(TEMPLATE)[0]
    o64, o80 = Oracle(20), Oracle(20, long_double=True)
    pos, vel = make_state(xyz, 20, o64.box, 7, spacing=8.6)
    o64.set_state(pos, vel)
    o80.set_state(pos, vel)
    for a, b in zip(o64.forces(), o80.forces()):
        assert rel_norm(a, b) < 1e-13
    e64, e80 = o64.energies(), o80.energies()
    np.testing.assert_allclose(e64[:5], e80[:5], rtol=1e-13)
    o64.run(50)
    o80.run(50)
    assert rel_norm(o64.get_state()[0], o80.get_state()[0]) < 1e-11


def _etot(o, with_thermostat):
    e = o.energies()
    return e[:5].sum() + (e[5] + e[6] if with_thermostat else 0.0)


def test_nve_energy_conservation():
    xyz = read_template_xyz(TEMPLATE)[0]
    o = Oracle(8, q_chain=1, C_uM=300.0, L=3.0, B=5)
    pos, vel = make_state(xyz, 8, o.box, 2, spacing=9.0, perturb=0.02)
    o.set_state(pos, vel)
    e0 = _etot(o, False)
    ke0 = o.energies()[0]
    es = []
    for _ in range(20):
        o.run(50)
        es.append(_etot(o, False))
    drift = max(abs(e - e0) for e in es)
    assert drift < 2e-3 * ke0, (drift, ke0)


def test_nose_hoover_extended_energy_conserved():
    xyz = read_template_xyz(TEMPLATE)[0]
    o = Oracle(8, q_chain=5, Q=10.0, K=350.0, C_uM=300.0, L=3.0, B=5)
    pos, vel = make_state(xyz, 8, o.box, 4, spacing=9.0, perturb=0.02, T=298.0 / 298.0)
    o.set_state(pos, vel)
    e0 = _etot(o, True)
    ke0 = o.energies()[0]
    es, xis = [], []
    for _ in range(20):
        o.run(50)
        es.append(_etot(o, True))
        xis.append(o.get_state()[2][0])
    assert max(abs(e - e0) for e in es) < 2e-3 * ke0
    # the thermostat is active: target T = 350/298 > initial T = 1 drives xi_0 negative
    assert min(xis) < -1e-3


def test_set_thermostat_continues_a_run():
    # a run stopped after 30 steps and continued in a fresh oracle from (pos, vel, xi, eta,
    # particle KE) follows the uninterrupted run to round-off; a restart that resets the chain
    # (the reference's restart) does not
    xyz = read_template_xyz(TEMPLATE)[0]
    kw = dict(q_chain=5, Q=10.0, K=350.0, C_uM=300.0, L=3.0, B=5)
    a = Oracle(8, **kw)
    pos, vel = make_state(xyz, 8, a.box, 5, spacing=9.0, perturb=0.02)
    a.set_state(pos, vel)
    a.run(30)
    p30, v30, xi, eta = a.get_state()
    assert np.all(np.abs(xi[:-1]) > 1e-6)  # the last link (Q = 0) stays at 0
    ke = a.energies()[0]
    a.run(30)
    pa, va, xia, _ = a.get_state()
    b = Oracle(8, **kw)
    b.set_state(p30, v30)
    b.set_thermostat(xi, eta, ke, step_index=30)
    b.run(30)
    pb, vb, xib, _ = b.get_state()
    assert rel_norm(vb, va) < 1e-10 and rel_norm(pb, pa) < 1e-12
    assert np.allclose(xib, xia, rtol=1e-9, atol=1e-14)
    c = Oracle(8, **kw)
    c.set_state(p30, v30)
    c.step_index = 30
    c.run(30)
    assert rel_norm(c.get_state()[1], va) > 1e-6
    with pytest.raises(RuntimeError):
        b.set_thermostat(xi[:3], eta[:3], ke)


def test_bd_fluctuation_dissipation(tmp_path):
    # free beads (one uncharged bead per subunit, far apart) starting at rest heat up as
    # <v_x^2> = (kT/m) (1 - exp(-2 t / damp)) with kT = 1 (src/md.cpp:415-498)
    path = write_template(str(tmp_path / "free"), [(0, 0, 0, 1, 0.0, 1.0, 1.0)], [], [], [])
    n = 1728
    damp, dt, steps = 10.0, 0.005, 600
    o = Oracle(n, template=path, C_uM=0.5, damp=damp, dt=dt, brownian=True, L=1.0, B=10 ** 6)
    g = np.arange(12) - 5.5
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * (o.box / 12)
    o.set_state(pos, np.zeros_like(pos))
    o.run(steps)
    v = o.get_state()[1]
    expect = 1.0 - math.exp(-2 * steps * dt / damp)
    got = float(np.mean(v * v))
    assert abs(got - expect) < 4 * expect * math.sqrt(2.0 / v.size), (got, expect)
    assert np.allclose(v.sum(axis=0), 0, atol=1e-10)


def _minstd_uniforms(skip, n):
    """Independent restatement: minstd_rand0 (x <- 16807 x mod 2^31-1, seed 69521) feeding
    generate_canonical<double, 53> (two draws, each double operation rounded) and
    uniform_real_distribution(-0.5, 0.5)."""
    M = 2147483647
    x = 69521
    R = 2147483646.0
    out = []
    total = skip + n
    for k in range(total):
        x = (x * 16807) % M
        g1 = x
        x = (x * 16807) % M
        g2 = x
        if k < skip:
            continue
        s = 0.0
        s += float(g1 - 1) * 1.0
        s += float(g2 - 1) * R
        u = s / (R * R)
        if u >= 1.0:
            u = math.nextafter(1.0, 0.0)
        out.append(u * 1.0 + (-0.5))
    return np.array(out)


def test_bd_stream_independent_restatement():
    for skip in [0, 12345]:
        a = bd_uniforms(skip, 4000)
        b = _minstd_uniforms(skip, 4000)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("L", [7.5, 0.3])
def test_reference_list_covers_force_cutoffs(L):
    xyz = read_template_xyz(TEMPLATE)[0]
    o = Oracle(20, L=L)
    pos, vel = make_state(xyz, 20, o.box, 12, spacing=8.4)
    o.set_state(pos, vel)
    q = np.tile(read_template_xyz(TEMPLATE)[2], 20)
    sub = np.repeat(np.arange(20), 105)
    box = o.box
    d = pos[:, None, :] - pos[None, :, :]
    d = np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))
    r = np.sqrt((d * d).sum(-1))
    need_lj = (sub[:, None] != sub[None, :]) & (r < o.ecut)
    need_q = (q[:, None] != 0) & (q[None, :] != 0) & (r < o.ecut_el)
    np.fill_diagonal(need_q, False)
    need = need_lj | need_q
    for i in range(len(pos)):
        nb = set(o.neighbours(i).tolist())
        missing = set(np.nonzero(need[i])[0].tolist()) - nb
        assert not missing, (i, sorted(missing)[:5])
