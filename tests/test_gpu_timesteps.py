"""GPU timestep criteria (agrav_gravity_timesteps) against the oracle's restatement of get_timestep_gravity
and find_long_range_step_constraint (timestep.c:284-344, :525-565) applied to the same step's outputs."""
import numpy as np
import pytest

from arepo_amd import ics

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["treepm", "isolated", "treepm_comoving"])
def test_timesteps_match_oracle(agrav, oracle, case):
    if case.startswith("treepm"):
        pos, mass, eps = ics.clustered_box(32768, nhalo=16, seed=12)
        kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(eps, 2 * eps))
    else:
        pos, mass, eps = ics.halo_disk(32768, seed=13)
        kw = dict(periodic=False, box_size=0.0, pmgrid=0, soft=(eps, 2 * eps))
    st = (np.arange(len(pos)) % 3 == 0).astype(np.uint8)
    mass = mass.copy()
    mass[5] = 0.0
    tkw = dict(eta=0.025, max_dt=0.05, min_dt=1e-9)
    if case.endswith("comoving"):
        tkw.update(time_a=0.5, hubble_a=2.0, comoving=True)
    with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP32, walk=agrav.WALK_GROUP)) as g:
        g.set_particles(pos, mass, softtype=st)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
        comp = g.get_components(tree=False, pm=kw["pmgrid"] > 0)
        dt, dd, below = g.gravity_timesteps(**tkw)
    p = oracle.make_params(**kw)
    odt, odd, obelow = oracle.gravity_timesteps(p, r["acc"], mass, acc_pm=comp["acc_pm"], stype=st, **tkw)
    assert dd == pytest.approx(odd, rel=1e-14)
    np.testing.assert_allclose(dt, odt, rtol=1e-14)
    assert below == obelow
    assert dt[5] == min(0.05, dd)
    if kw["pmgrid"] == 0:
        assert dd == 0.05


def test_timesteps_need_a_step(agrav):
    pos, mass, eps = ics.uniform_box(4096, seed=3)
    with agrav.Gravity(agrav.make_params(periodic=True, box_size=1.0, pmgrid=16, soft=(eps,))) as g:
        g.set_particles(pos, mass)
        with pytest.raises(agrav.AgravError) as ei:
            g.gravity_timesteps()
        assert ei.value.code == agrav.ERR_STATE
