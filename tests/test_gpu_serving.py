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
    fp = agrav.solve_footprint(_params(agrav), 20000)
    budget = int(1.6 * fp)
    # budget below two requests' footprint: the server must run them one at a time, and still finish all
    with agrav.Server(_params(agrav), slots=4, hbm_budget_bytes=budget) as s:
        tickets = [s.submit(p, m) for p, m in sets]
        outs = [s.wait(t) for t in tickets]
        st = s.stats()
    print("footprint %d, budget %d, stats %s" % (fp, budget, st))
    assert st["peak_running"] == 1 and st["completed"] == 6
    assert st["hbm_peak"] <= budget
    assert st["model_exceeded"] == 0
    with agrav.Gravity(_params(agrav)) as g:
        g.set_particles(*sets[3])
        g.compute_grav_accelerations(first_step=True)
        ref = g.get_accelerations()
    assert np.array_equal(outs[3]["acc"], ref["acc"])


def test_resident_buffers_are_charged_and_evicted(agrav):
    """Slots keep their buffers between requests; admission charges them, and an idle
    slot's buffers are dropped when only they keep a request out."""
    big = _hernquist(150000, 210)
    small = [_hernquist(3000, 211 + i) for i in range(3)]
    p = _params(agrav)
    budget = int(1.3 * agrav.solve_footprint(p, 150000))
    with agrav.Server(p, slots=2, hbm_budget_bytes=budget) as s:
        s.wait(s.submit(*big))
        st = s.stats()
        assert st["hbm_resident"] > agrav.solve_footprint(p, 3000)  # the big solve's buffers stay resident
        # a second big request on the other slot cannot fit beside the first slot's resident buffers
        t = [s.submit(*big)] + [s.submit(pp, m) for pp, m in small]
        outs = [s.wait(x) for x in t]
        st = s.stats()
    # whichever slot takes the second big request, the charged total stays within the budget (the slot
    # that did not solve the first one can only take it after evicting the other's buffers)
    assert st["hbm_peak"] <= budget and st["model_exceeded"] == 0
    assert st["completed"] == 5 and st["failed"] == 0
    with agrav.Gravity(p) as g:
        g.set_particles(*big)
        g.compute_grav_accelerations(first_step=True)
        ref = g.get_accelerations()
    assert np.array_equal(outs[0]["acc"], ref["acc"])


@pytest.mark.parametrize("n", [10_000, 100_000, 1_000_000])
def test_footprint_model_bounds_device_memory(agrav, n):
    """The footprint model bounds what a context holds after a solve: the sum of its
    device buffers, and the drop in the device's free memory (hipMemGetInfo)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipDeviceSynchronize() == 0 and hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    p = _params(agrav)
    pos, mass = _hernquist(n, 220)
    before = free_bytes()
    with agrav.Gravity(p) as g:
        base = g.timings()["device_bytes"]
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        g.get_accelerations()
        held = g.timings()["device_bytes"] - base
        drop = before - free_bytes()
    model = agrav.solve_footprint(p, n)
    print("n=%d: buffers %.1f MB, free-memory drop %.1f MB, model %.1f MB" % (n, held / 1e6, drop / 1e6, model / 1e6))
    assert held <= model
    assert drop <= model + base + (8 << 20)


def test_request_softening_zero_uses_server_table(agrav):
    """softening = 0 means the server's parameter, whatever the slot solved before."""
    a, b = _hernquist(4000, 230), _hernquist(4000, 231)
    with agrav.Server(_params(agrav), slots=1) as s:
        s.wait(s.submit(*a, softening=0.03))
        out = s.wait(s.submit(*b, softening=0.0))
    with agrav.Gravity(_params(agrav)) as g:
        g.set_particles(*b)
        g.compute_grav_accelerations(first_step=True)
        ref = g.get_accelerations()
    assert np.array_equal(out["acc"], ref["acc"]) and np.array_equal(out["pot"], ref["pot"])


def test_destroy_with_queued_requests_returns(agrav):
    """Closing the server with requests still queued under a tight budget finishes them and returns."""
    sets = [_hernquist(20000, 240 + i) for i in range(5)]
    p = _params(agrav)
    s = agrav.Server(p, slots=3, hbm_budget_bytes=int(1.2 * agrav.solve_footprint(p, 20000)))
    for pp, m in sets:
        s.submit(pp, m)
    s.close()  # must not hang


def test_request_beyond_budget_is_rejected(agrav):
    p = _params(agrav)
    with agrav.Server(p, slots=1, hbm_budget_bytes=int(agrav.solve_footprint(p, 1000) * 1.5)) as s:
        pos, mass = _hernquist(100000, 250)
        with pytest.raises(agrav.AgravError) as ei:
            s.submit(pos, mass)
        assert ei.value.code == agrav.ERR_CAPACITY


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


def test_request_stream_at_full_rate(agrav):
    """The serving workload's request stream (Poisson arrivals at 200 per second, isolated Hernquist sets
    with N log-normal around 5e4, sigma 1.2 in ln N, in [256, 4M]) for 8 s in real time on 4 slots: every
    request completes, the server keeps up with the arrivals, stays within its HBM budget, and sampled
    results equal solves of the same sets alone."""
    import time
    from arepo_amd import ics
    g = ics.rng(71)
    pool_n = 1 << 22
    pool = ics._isotropic(ics.hernquist_radii(pool_n, 1.0, g), g)
    rate, seconds = 200.0, 8.0
    t_arr = np.cumsum(g.exponential(1.0 / rate, int(rate * seconds * 1.5) + 10))
    t_arr = t_arr[t_arr < seconds]
    sizes = np.clip(np.exp(np.log(5e4) + 1.2 * g.standard_normal(len(t_arr))), 256, 4e6).astype(np.int64)
    starts = g.integers(0, pool_n - sizes + 1)
    eps = (1.0 + np.sqrt(2.0)) / 50.0
    p = agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP, group_size=0, leaf_bucket=4, eval_potential=True)
    budget = 64 << 30
    sample = set(range(0, len(t_arr), 200))
    kept = {}
    with agrav.Server(p, slots=4, hbm_budget_bytes=budget) as srv:
        tickets = []
        t0 = time.perf_counter()
        for k, (ta, n, s0) in enumerate(zip(t_arr, sizes, starts)):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            tickets.append(srv.submit(pool[s0:s0 + n], np.full(n, 1.0 / n)))
        outs = []
        for k, t in enumerate(tickets):
            r = srv.wait(t)
            if k in sample:
                kept[k] = r["acc"].copy()
            outs.append(r["total_ms"])
        makespan = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array(outs)
    print("%d requests (mean N %.3g) in %.2f s: latency p50 %.2f ms, p99 %.2f ms; %s"
          % (len(tickets), sizes.mean(), makespan, np.percentile(lat, 50), np.percentile(lat, 99), st))
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert makespan < seconds + 2.0  # kept up with the arrivals
    assert st["hbm_peak"] <= budget and st["model_exceeded"] == 0
    with agrav.Gravity(p) as g1:
        for k, acc in kept.items():
            n, s0 = int(sizes[k]), int(starts[k])
            g1.set_particles(pool[s0:s0 + n], np.full(n, 1.0 / n))
            g1.compute_grav_accelerations(first_step=True)
            assert np.array_equal(g1.get_accelerations(fields=("acc",))["acc"], acc), k
