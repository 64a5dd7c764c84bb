"""Batched parameter-sweep replicas, the Verlet skin and gather tile, and the autotuner.

Forces are summed per bead in list order over lists that hold every pair inside the cutoffs,
so the skin, the tile, the tuner's trials and the batching must leave every trajectory
unchanged bit for bit.
"""
import numpy as np
import pytest

import souffle_gpu as sg
import replica_sweep
from oracle_lib import TEMPLATE

pytestmark = pytest.mark.gpu


def sweep(n, seed, ns_range=(20, 60), fp32=True):
    t = sg.Template(TEMPLATE)
    sizes, params, pos, vel, _ = replica_sweep.sweep(t, n, seed, ns_range=ns_range, fp32=fp32)
    return t, sizes, params, pos, vel


def test_skin_and_tile_leave_trajectory_unchanged():
    t, sizes, params, pos, vel = sweep(1, 3, (30, 30), fp32=False)
    a = sg.System(t, sizes, params)
    a.set_state(pos[0], vel[0])
    a.step(150)
    ap, av = a.get_state()
    b = sg.System(t, sizes, params)
    b.set_state(pos[0], vel[0])
    b.step(50)
    b.set_skin(1.0)
    b.set_tile(128)
    b.step(50)
    assert b.skin() == 1.0
    b.set_skin(0.15)
    b.set_tile(512)
    b.step(25)
    b.set_tile(256)  # the first 50 steps ran at the default tile, 384
    b.step(25)
    bp, bv = b.get_state()
    assert np.array_equal(ap, bp) and np.array_equal(av, bv)
    with pytest.raises(sg.SfError):
        b.set_tile(64)
    with pytest.raises(sg.SfError):
        b.set_skin(params[0].box)  # cutoff + skin beyond half a box


def test_autotune_keeps_state_and_trajectory():
    t, sizes, params, pos, vel = sweep(3, 5)
    x0, v0 = np.concatenate(pos), np.concatenate(vel)
    a = sg.System(t, sizes, params)
    a.set_state(x0, v0)
    a.step(120)
    ap, av = a.get_state()
    b = sg.System(t, sizes, params)
    b.set_state(x0, v0)
    res, ms = b.autotune(skins=[0.3, 0.8], tiles=[128, 512], trial_steps=40)
    assert res["n_trials"] == 4 and ms.shape == (2, 2) and (ms > 0).all()
    assert res["skin"] in (0.3, 0.8) and res["tile"] in (128, 512)
    assert abs(res["ms_per_step"] - ms.min()) < 1e-12
    assert b.stats()["steps"] == 0
    p0, v0b = b.get_state()
    assert np.array_equal(p0, x0) and np.array_equal(v0b, v0)
    b.step(120)
    bp, bv = b.get_state()
    assert np.array_equal(ap, bp) and np.array_equal(av, bv)
    assert b.stats()["steps"] == 120


def test_sweep_batch_matches_single_runs():
    """Sixteen uneven replicas of a salt / epsilon / temperature sweep in one handle (FP32 pair
    arithmetic, autotuned) against the same replicas run one by one."""
    t, sizes, params, pos, vel = sweep(16, 11)
    batch = sg.System(t, sizes, params)
    batch.set_state(np.concatenate(pos), np.concatenate(vel))
    batch.autotune(skins=[0.3, 0.6], tiles=[256], trial_steps=30)
    batch.step(200)
    bp, bv = batch.get_state()
    assert np.isfinite(bp).all()
    off = 0
    for k, ns in enumerate(sizes):
        single = sg.System(t, ns, params[k])
        single.set_state(pos[k], vel[k])
        single.step(200)
        sp, sv = single.get_state()
        n = ns * t.n_beads
        assert np.array_equal(bp[off:off + n], sp), k
        assert np.array_equal(bv[off:off + n], sv), k
        # each replica keeps its own thermostat
        xb, _ = batch.thermostat(k)
        xs, _ = single.thermostat(0)
        assert np.array_equal(xb, xs), k
        off += n


def test_whole_run_in_one_graph_launch():
    """A sweep batch run as one HIP graph launch of all its steps (no host round trip) against the
    same batch in 100-step launches: the same trajectories bit for bit."""
    t, sizes, params, pos, vel = sweep(12, 17)
    x0, v0 = np.concatenate(pos), np.concatenate(vel)
    a = sg.System(t, sizes, params)
    a.set_state(x0, v0)
    a.step(350)
    st_a = a.stats()
    ap, av = a.get_state()
    b = sg.System(t, sizes, params)
    b.set_state(x0, v0)
    b.set_graph_steps(350)
    n0 = b.stats()["graph_launches"]
    b.step(350)
    st_b = b.stats()
    bp, bv = b.get_state()
    assert st_b["graph_launches"] - n0 == 1 and st_b["graph_steps"] == 350
    assert st_a["graph_steps"] == 100
    assert np.array_equal(ap, bp) and np.array_equal(av, bv)
    assert st_a["steps"] == st_b["steps"] == 350 and st_a["rebuilds"] == st_b["rebuilds"]
    for k in range(len(sizes)):
        assert np.array_equal(a.thermostat(k)[0], b.thermostat(k)[0]), k
    with pytest.raises(sg.SfError):
        b.set_graph_steps(20001)


def test_long_graph_replays_after_list_growth():
    """A larger skin set before a 150-step launch makes its first list build outgrow the
    capacities: the launch is replayed from its start with grown lists, and the run equals the
    one in 100-step launches."""
    from souffle_states import make_state, read_template_xyz
    t = sg.Template(TEMPLATE)
    p = t.params(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, n_subunits=27)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], 27, p.box, 11, spacing=7.6, perturb=0.02)
    runs = []
    for steps in (0, 150):
        s = sg.System(t, 27, p)
        s.set_graph_steps(steps)
        s.set_state(pos, vel)
        s.set_skin(2.5)
        st0 = s.stats()
        s.step(150)
        runs.append((s.get_state(), st0, s.stats()))
    (ap, av), _, sa = runs[0]
    (bp, bv), b0, sb = runs[1]
    print("capacities", b0["lj_capacity"], b0["q_capacity"], "->", sb["lj_capacity"], sb["q_capacity"],
          "launches", sb["graph_launches"] - b0["graph_launches"], flush=True)
    assert np.array_equal(ap, bp) and np.array_equal(av, bv)
    assert sb["lj_capacity"] == sa["lj_capacity"] and sb["q_capacity"] == sa["q_capacity"]
    assert (sb["lj_capacity"], sb["q_capacity"]) != (b0["lj_capacity"], b0["q_capacity"])  # the lists grew
    assert sb["graph_launches"] - b0["graph_launches"] >= 2  # the overflowed launch and its replay


def test_long_graph_stops_at_the_last_finite_step(tmp_path):
    """A run that goes non-finite inside one 50-step launch is replayed up to its last finite step."""
    from template_writer import write_template
    path = write_template(str(tmp_path / "point"), beads=[(0.0, 0.0, 0.0, 1, 0.0, 0.01, 1.0)])
    t = sg.Template(path)
    p = t.params(n_subunits=2, C_uM=50.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, q_chain=1)
    s = sg.System(t, 2, p)
    s.set_graph_steps(50)
    s.set_state(np.array([[-5.0, 0.0, 0.0], [5.0, 0.0, 0.0]]), np.array([[100.0, 0.0, 0.0], [-100.0, 0.0, 0.0]]))
    with pytest.raises(sg.SfError) as e:
        s.step(50)
    assert e.value.code == sg.SF_E_NONFINITE
    gp, _ = s.get_state()
    assert np.array_equal(gp[:, 0], [-0.5, 0.5]) and s.stats()["steps"] == 9
