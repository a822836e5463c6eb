"""GPU tests of the continuous-batching server (agrav_server_*, through the C ABI).

The key property: a request's result is bitwise identical whether it runs alone
or in a mix of concurrent requests (every kernel is deterministic and each slot
owns its buffers and stream).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _hernquist(n, seed, a=1.0):
    from arepo_amd import ics
    g = ics.rng(seed)
    pos = ics._isotropic(ics.hernquist_radii(n, a, g), g)
    return pos, np.full(n, 1.0 / n)


def _params(agrav, precision=None):
    return agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.05,),
                             precision=agrav.FP32 if precision is None else precision, walk=agrav.WALK_GROUP)


def test_result_same_alone_and_in_mix(agrav):
    sets = [_hernquist(n, 100 + i) for i, n in enumerate([5000, 300, 40000, 1200, 20000, 7000, 256, 90000])]
    target = 2
    with agrav.Server(_params(agrav), slots=1) as s:
        alone = s.wait(s.submit(*sets[target], softening=0.03))
    with agrav.Server(_params(agrav), slots=4) as s:
        tickets = [s.submit(p, m, softening=0.03 if i == target else 0.05) for i, (p, m) in enumerate(sets)]
        # a second copy of the target request, submitted while the others run
        tickets.append(s.submit(*sets[target], softening=0.03))
        outs = [s.wait(t) for t in tickets]
        st = s.stats()
    assert np.array_equal(outs[target]["acc"], alone["acc"])
    assert np.array_equal(outs[target]["pot"], alone["pot"])
    assert np.array_equal(outs[-1]["acc"], alone["acc"])
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["peak_running"] >= 2  # requests did overlap


def test_admission_by_footprint(agrav):
    sets = [_hernquist(20000, 200 + i) for i in range(6)]
    # budget below two requests' footprint: the server must run them one at a time, and still finish all
    with agrav.Server(_params(agrav), slots=4, hbm_budget_bytes=100 << 20) as s:
        tickets = [s.submit(p, m) for p, m in sets]
        outs = [s.wait(t) for t in tickets]
        st = s.stats()
    assert st["peak_running"] == 1 and st["completed"] == 6
    assert st["hbm_peak"] <= 100 << 20 or st["peak_running"] == 1
    with agrav.Gravity(_params(agrav)) as g:
        g.set_particles(*sets[3])
        g.compute_grav_accelerations(first_step=True)
        ref = g.get_accelerations()
    assert np.array_equal(outs[3]["acc"], ref["acc"])


def test_server_results_are_accurate(agrav, oracle):
    pos, mass = _hernquist(3000, 300)
    with agrav.Server(_params(agrav), slots=2) as s:
        out = s.wait(s.submit(pos, mass, softening=0.05))
    op = oracle.make_params(periodic=False, box_size=0.0, soft=(0.05,))
    ad, pd = oracle.direct_summation(op, pos, mass)
    e = rel_err(out["acc"], ad)
    assert np.median(e) < 2.5e-3 and np.percentile(e, 99) < 2e-2


def test_server_rejects_unknown_ticket(agrav):
    with agrav.Server(_params(agrav), slots=1) as s:
        with pytest.raises(agrav.AgravError):
            s.wait(123456)
