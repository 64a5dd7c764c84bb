"""Pins the oracle's encapsulation big beads against what libc and the mathematics fix.

Reference: placement src/initialize.cpp:566-680 (srand(0), rand() triples, overlap rejection),
forces src/forces.cpp:174-236, energies src/energies.cpp:96-136, the big beads' BD draws
src/md.cpp:482-487, defaults src/md.cpp:86-90 (radius 3.0, so diameter 6; charge -78;
hard-core distance 3.5).

- placement: the first candidates of the C library's rand() stream (read through ctypes here)
  that keep clear of every subunit bead and earlier big bead, recomputed in float64;
- forces: minus the finite-difference gradient of the big-bead energy (sum of ce and ne), for
  beads inside the shifted-WCA range and beads reached by the Yukawa term only;
- energies: the subunit totals ne and ce grow by exactly the big beads' totals (each term is
  added at full value to the subunit bead), the other totals do not move;
- BD: the big beads' draws come after the subunit beads' draws of a step (src/md.cpp:482-487),
  so a non-interacting big bead leaves the first step unchanged and changes the second.
  The number of draws (three per big bead) rests on those source lines only.
"""

import ctypes

import numpy as np

from oracle_lib import TEMPLATE, Oracle, place_big_beads, rel_norm
from souffle_states import make_state, read_template_xyz

BIG_SIGMA, BIG_Q, SHC = 6.0, -78.0, 3.5
RAND_MAX = 2147483647


def _system(seed=3, long_double=False, brownian=False):
    o = Oracle(20, long_double=long_double, brownian=brownian)
    xyz = read_template_xyz(TEMPLATE)[0]
    pos, vel = make_state(xyz, 20, o.box, seed, spacing=8.6, perturb=0.05)
    return o, pos, vel


def _wrap(d, box):
    return d - box * np.floor(d / box + 0.5)


def test_placement_is_first_clear_candidate_of_rand_stream():
    o, pos, _ = _system()
    sig, _ = o.bead_props()
    n = 3
    got = place_big_beads(pos, sig, o.box, n, BIG_SIGMA)
    assert got is not None
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(0)
    placed, tried = [], 0
    for _ in range(n):
        while True:
            c = (np.array([libc.rand(), libc.rand(), libc.rand()], dtype=np.float64) / RAND_MAX - 0.5) * o.box
            tried += 1
            d = _wrap(c - pos, o.box)
            ok = bool(np.all(np.sum(d * d, axis=1) >= ((sig + BIG_SIGMA) / 2) ** 2))
            if ok and placed:
                dd = _wrap(c - np.array(placed), o.box)
                ok = bool(np.all(np.sum(dd * dd, axis=1) >= BIG_SIGMA ** 2))
            if ok:
                placed.append(c)
                break
    assert tried > n, "every first candidate was clear; the rejection path is untested"
    np.testing.assert_allclose(got, np.array(placed), rtol=0, atol=1e-12)
    # a bead that needs more tries than allowed is an error, as the reference's exit(1)
    assert place_big_beads(pos, sig, o.box, n, BIG_SIGMA, max_attempts=1) is None


def _big_position(o, pos, sig):
    """A big bead 3.7 from one bead (inside the WCA range 3.5-3.93 of a sigma-1 bead)."""
    c = place_big_beads(pos, sig, o.box, 1, BIG_SIGMA)[0]
    d = _wrap(pos - c, o.box)
    r = np.linalg.norm(d, axis=1)
    j = int(np.argmin(r))
    return _wrap(c + d[j] * (1.0 - 3.7 / r[j]), o.box)


def _big_energy(o, pos, vel):
    o.set_state(pos, vel)
    o.energies()
    return float(np.sum(o.big_energies()))


def test_big_bead_force_is_minus_energy_gradient():
    o, pos, vel = _system()
    sig, q = o.bead_props()
    c = _big_position(o, pos, sig)
    o.set_state(pos, vel)
    F0 = o.forces()[0]
    o.set_big_beads(c, BIG_SIGMA, BIG_Q, SHC)
    F1 = o.forces()[0]
    Fbig = F1 - F0
    r = np.linalg.norm(_wrap(c - pos, o.box), axis=1)
    delta = (BIG_SIGMA + sig) / 2 - SHC
    wca = r < delta + 1.12246205 * SHC
    yk = (q != 0) & (r < o.ecut_el)
    assert wca.sum() >= 1 and (yk & ~wca).sum() >= 1
    # beads away from either cutoff, several of each kind
    far_from_cut = (np.abs(r - (delta + 1.12246205 * SHC)) > 1e-3) & (np.abs(r - o.ecut_el) > 1e-3)
    pick = list(np.flatnonzero(wca & far_from_cut)[:3]) + list(np.flatnonzero(yk & ~wca & far_from_cut)[:3])
    h = 1e-5
    for j in pick:
        g = np.zeros(3)
        for k in range(3):
            p = pos.copy()
            p[j, k] += h
            ep = _big_energy(o, p, vel)
            p[j, k] -= 2 * h
            em = _big_energy(o, p, vel)
            g[k] = (ep - em) / (2 * h)
        assert rel_norm(Fbig[j], -g) < 1e-6, (j, Fbig[j], -g)
    # beads out of reach feel nothing from the big bead
    out = ~(wca | yk)
    assert out.sum() > 0
    np.testing.assert_array_equal(F1[out], F0[out])


def test_big_bead_energy_terms_go_to_subunit_totals():
    o, pos, vel = _system()
    sig, _ = o.bead_props()
    c = _big_position(o, pos, sig)
    o.set_state(pos, vel)
    t0 = o.energies()
    o.set_big_beads(c, BIG_SIGMA, BIG_Q, SHC)
    t1 = o.energies()
    ce_big, ne_big = o.big_energies()
    assert ce_big != 0 and ne_big > 0
    # totals: ke, se, be, ne, ce, tpe, tke
    for k in (0, 1, 2, 5, 6):
        assert t1[k] == t0[k]
    assert abs((t1[3] - t0[3]) - ne_big) <= 1e-12 * max(1.0, abs(t1[3]))
    assert abs((t1[4] - t0[4]) - ce_big) <= 1e-12 * max(1.0, abs(t1[4]))


def test_long_double_and_double_big_bead_forces_agree():
    o, pos, vel = _system()
    sig, _ = o.bead_props()
    c = _big_position(o, pos, sig)
    F = []
    for ld in (False, True):
        oo, _, _ = _system(long_double=ld)
        oo.set_state(pos, vel)
        oo.set_big_beads(c, BIG_SIGMA, BIG_Q, SHC)
        F.append(oo.forces()[0])
    assert rel_norm(F[0], F[1]) < 1e-13


def test_big_bead_draws_follow_the_subunit_draws():
    o, pos, vel = _system(brownian=True)
    sig, _ = o.bead_props()
    c = place_big_beads(pos, sig, o.box, 1, BIG_SIGMA)[0]
    runs = []
    for with_big in (False, True):
        oo, _, _ = _system(brownian=True)
        if with_big:
            oo.set_big_beads(c, 0.01, 0.0, 0.01)  # uncharged and tiny: no force on any bead
        oo.set_state(pos, vel)
        oo.run(1)
        v1 = oo.get_state()[1]
        oo.run(1)
        v2 = oo.get_state()[1]
        runs.append((v1, v2))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert not np.array_equal(runs[0][1], runs[1][1])
