"""sf_oligomers_noalloc (the allocation-free oligomer entry point) against sf_oligomers through
the C ABI: labels, mass spectra and the oligomer count must be identical (integers, bit-exact)
on consecutive calls that share the GPU's fixed buffers, the second on a smaller handle, and a
replica batch (per-replica slots of the buffers). Handles beyond the buffers' size are refused
with a clear error before any device work."""
import numpy as np
import pytest

import souffle_gpu as sg
from oracle_lib import TEMPLATE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def template():
    return sg.Template(TEMPLATE)


def _evolved(template, ns, seed, steps=300, C_uM=4000.0):
    """A seeded state at high concentration, stepped so that subunits touch."""
    p = template.params(n_subunits=ns, C_uM=C_uM, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
    pos, vel = template.make_state(ns, p.box, p.T, seed)
    s = sg.System(template, ns, p)
    s.set_state(pos, vel)
    s.step(steps)
    return s


def _same(s, contact):
    ref = s.oligomers(contact)
    got, n = s.oligomers_noalloc(contact)
    assert len(ref) == len(got)
    for (lr, mr), (lg, mg) in zip(ref, got):
        assert np.array_equal(lr, lg)
        assert np.array_equal(mr, mg)
    assert n == sum(int(m.sum()) for _, m in ref)
    return ref, n


def test_noalloc_matches_sf_oligomers_twice(template):
    big = _evolved(template, 300, 11)
    ref, n_big = _same(big, 2.5)
    assert n_big < 300, "the state has no contacts: the comparison would be vacuous"
    small = _evolved(template, 40, 12, steps=100)
    _, n_small = _same(small, 2.5)
    assert n_small <= 40
    # the larger handle again, after the smaller one left its contents in the buffers
    _same(big, 2.5)
    big.close()
    small.close()


def test_noalloc_replica_batch(template):
    sizes = [40, 60, 25]  # at 2000 uM the smallest box still holds the Yukawa cutoff + skin
    params, pos, vel = [], [], []
    for k, ns in enumerate(sizes):
        p = template.params(n_subunits=ns, C_uM=2000.0, c_mM=[100.0, 300.0, 150.0][k], elj_att=2.1, dt=0.005,
                            K=298.0)
        x, v = template.make_state(ns, p.box, p.T, 40 + k)
        params.append(p)
        pos.append(x)
        vel.append(v)
    s = sg.System(template, sizes, params)
    s.set_state(np.concatenate(pos), np.concatenate(vel))
    s.step(200)
    _same(s, 2.5)
    s.close()


def test_noalloc_refuses_handles_beyond_its_buffers(template):
    ns = (1 << 18) + 1
    p = template.params(n_subunits=ns, C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0)
    s = sg.System(template, ns, p)
    with pytest.raises(sg.SfError, match="at most 262144 subunits"):
        s.oligomers_noalloc(2.0)
    s.close()
