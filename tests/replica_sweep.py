"""Seeded parameter sweep of independent replicas: uneven subunit counts, and salt, LJ well
depth and temperature drawn per replica; one seeded random state per replica (the library's
sf_make_state: non-overlapping subunits, random orientations, Maxwell-Boltzmann velocities).
Shared by the replica tests and the benchmark's replica workload."""
import numpy as np

import souffle_gpu as sg
from souffle_states import read_template_xyz


def sweep(t, n, seed, ns_range=(20, 400), c_range=(50.0, 800.0), eps_range=(1.5, 2.5),
          K_range=(280.0, 320.0), fp32=True, C_uM=800.0, dt=0.005, q_chain=5, Q=10.0):
    """Returns (sizes, params, pos, vel, draws) for n replicas of template t."""
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(ns_range[0], ns_range[1] + 1, n)]
    params, pos, vel, draws = [], [], [], []
    for k, ns in enumerate(sizes):
        d = dict(c_mM=float(rng.uniform(*c_range)), elj_att=float(rng.uniform(*eps_range)),
                 K=float(rng.uniform(*K_range)))
        p = t.params(n_subunits=ns, C_uM=C_uM, dt=dt, q_chain=q_chain, Q=Q, **d)
        p.pair_fp32 = 1 if fp32 else 0
        x, v = t.make_state(ns, p.box, p.T, 1000 + 17 * k + seed)
        params.append(p)
        pos.append(x)
        vel.append(v)
        draws.append(d)
    return sizes, params, pos, vel, draws


def reference_list_radius(template_path, ecut, ecut_el):
    """Smallest -L (plus 0.5 sigma of slack for deformation) at which the reference's list
    keeps every pair inside the cutoffs: its subunit prefilter admits centre-bead distances
    below 1.5 (max(ecut, ecut_el) + L) (src/functions.cpp:82-91), and a bead pair inside a
    cutoff has centre distance up to cutoff + twice the bead's reach from the centre bead."""
    xyz, _, q, _ = read_template_xyz(template_path)
    c = np.argmin(np.linalg.norm(xyz - xyz.mean(axis=0), axis=1))
    reach = np.linalg.norm(xyz - xyz[c], axis=1)
    r_all = reach.max()
    r_q = reach[q != 0].max() if (q != 0).any() else 0.0
    need = max(ecut + 2 * r_all, ecut_el + 2 * r_q) + 0.5
    return max(0.3, need / 1.5 - max(ecut, ecut_el))
