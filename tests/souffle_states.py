"""Seeded test states for the SOUFFLE force-and-integrate path.

Test helper only: builds positions/velocities in numpy so the oracle and the GPU library
are fed identical inputs. Subunits are copies of the template, rigidly rotated by a
uniform random rotation and placed on a jittered cubic lattice, then every bead is
perturbed (bonded forces vanish at the undeformed template, so parity states must be
deformed). Velocities are Maxwell-Boltzmann at temperature T, centre-of-mass removed.
"""
import numpy as np


def read_template_xyz(path):
    """Bead coordinates, charges and masses of a template (format of the reference's
    infiles: three header lines, then `# Number of Particles = n`, a header line and
    `index x y z type charge diameter mass` rows)."""
    tok = open(path).read().split()
    np_ = int(tok[14])
    rows = tok[23:23 + 8 * np_]
    arr = np.array(rows, dtype=float).reshape(np_, 8)
    return arr[:, 1:4].copy(), arr[:, 4].astype(int), arr[:, 5].copy(), arr[:, 7].copy()


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def make_state(tmpl_xyz, n_sub, box, seed, spacing=None, jitter=0.5, perturb=0.05, T=1.0,
               masses=None, wrap=True, min_gap=0.9):
    """Return (pos, vel) as [N,3] float64 arrays, N = n_sub * n_template_beads. Each subunit's
    rotation is redrawn until its beads are at least `min_gap` from all beads placed before."""
    rng = np.random.default_rng(seed)
    np_ = tmpl_xyz.shape[0]
    com = tmpl_xyz.mean(axis=0)
    body = tmpl_xyz - com
    nfill = int(np.ceil(n_sub ** (1.0 / 3.0) - 1e-9))
    if spacing is None:
        spacing = box / nfill
    pos = np.empty((n_sub * np_, 3))
    k = 0
    for i in range(nfill):
        for j in range(nfill):
            for l in range(nfill):
                if k >= n_sub:
                    break
                c = (np.array([i, j, l]) + 0.5 - nfill / 2.0) * spacing + rng.uniform(-jitter, jitter, 3)
                for attempt in range(2000):
                    R = random_rotation(rng)
                    trial = body @ R.T + c
                    if k == 0:
                        break
                    d = trial[:, None, :] - pos[None, :k * np_, :]
                    d = np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))
                    if np.min(np.einsum("ijk,ijk->ij", d, d)) >= min_gap * min_gap:
                        break
                else:
                    raise RuntimeError("could not place subunits without overlap; increase spacing")
                pos[k * np_:(k + 1) * np_] = trial
                k += 1
    pos += rng.normal(scale=perturb, size=pos.shape)
    if wrap:
        pos = np.where(pos > box / 2, pos - box, pos)
        pos = np.where(pos < -box / 2, pos + box, pos)
    m = np.ones(len(pos)) if masses is None else np.tile(masses, n_sub)
    vel = rng.normal(size=pos.shape) * np.sqrt(T / m)[:, None]
    vel -= vel.mean(axis=0)
    ke = 0.5 * np.sum(m[:, None] * vel * vel)
    vel *= np.sqrt(1.5 * len(pos) * T / ke)
    return pos, vel


def _min_image(d, box):
    return np.where(d > box / 2, d - box, np.where(d < -box / 2, d + box, d))


def min_bead_distance(a, b, box):
    """Smallest single-shift minimum-image distance between the beads of a and of b."""
    d = _min_image(a[:, None, :] - b[None, :, :], box)
    return float(np.sqrt(np.einsum("ijk,ijk->ij", d, d).min()))


def subunit_line(tmpl_xyz, links, box, seed, touch=1.5, apart=2.5, origin=None):
    """Subunits in a row along x, randomly rotated: subunit k + 1 slides towards subunit k until
    their closest beads are `touch` apart when links[k] is true, `apart` otherwise. Returns
    positions [(len(links) + 1) * n_template_beads, 3], wrapped into the box."""
    rng = np.random.default_rng(seed)
    body = tmpl_xyz - tmpl_xyz.mean(axis=0)
    x0 = np.zeros(3) if origin is None else np.asarray(origin, dtype=float)
    subs = [body @ random_rotation(rng).T + x0]
    for link in links:
        nxt = body @ random_rotation(rng).T
        target = touch if link else apart
        xs = subs[-1][:, 0].max() - nxt[:, 0].min() + target + 0.5
        while True:
            trial = nxt + subs[-1].mean(axis=0) * [0, 1, 1] + [xs, 0, 0]
            if min_bead_distance(subs[-1], trial, 1e9) < target:
                break
            xs -= 0.005
        subs.append(trial)
    pos = np.concatenate(subs)
    return _min_image(pos, box)
