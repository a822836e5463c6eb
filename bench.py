#!/usr/bin/env python3
"""Benchmark of the TreePM gravity step on one MI355X.

Workload (the metric's configuration that fits one GPU): N = 128^3 = 2,097,152
particles in the periodic unit box, half uniform and half in 256 Hernquist
spheres, PMGRID = 128, ASMTH = 1.25, RCUT = 4.5, ErrTolForceAcc = 0.0025
(relative criterion), softening 1/40 of the mean spacing.

A step is one full gravity solve as AREPO does it every step
(compute_grav_accelerations with OldAcc from the previous step): periodic PM,
octree build, tree walk, finalize.  Inputs are resident in HBM before the
timed region.  `value` is particles per second of the FP32 group-walk path.

Also reported: interactions/s, the plain FP64 GPU path (one lane per target,
AREPO's per-particle criterion) and the FP64 group walk on the same inputs,
the relative force error of each against an FP64 direct (Ewald) summation on
a seeded random subset, the roofline of the dominant kernel (the walk), and
the CPU oracle timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s & interactions/s, 128^3 TreePM, FP32 vs FP64; p50/p99 latency"
HBM_PEAK = 8.0e12    # MI355X HBM3E, bytes/s
FP32_PEAK = 157.3e12  # MI355X FP32, flop/s


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def make_params(A, precision, walk, eps, pmgrid=128, pot=False, group_size=64):
    return A.make_params(periodic=True, box_size=1.0, pmgrid=pmgrid, asmth=1.25, rcut=4.5, G=1.0,
                         err_tol_theta=0.7, err_tol_force_acc=0.0025, opening=1, soft=(eps,),
                         precision=precision, walk=walk, group_size=group_size, leaf_bucket=4, eval_potential=pot)


def run_path(A, pos, mass, eps, precision, walk, steps, warmup, old_acc=None, pot=False, components=False):
    """Warm up, then time `steps` full solves; returns timing dict and the last result (with the tree and
    PM parts of the acceleration when `components`)."""
    g = A.Gravity(make_params(A, precision, walk, eps, pot=pot))
    g.set_particles(pos, mass)
    if old_acc is not None:
        g.set_old_acc(old_acc)
    first = old_acc is None
    for w in range(max(warmup, 1 if first else 0)):
        g.compute_grav_accelerations(first_step=first)
        first = False
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        g.compute_grav_accelerations(first_step=False)  # synchronises at the end (event wait)
        step_ms.append(1e3 * (time.perf_counter() - ts))
    total = time.perf_counter() - t0
    tim = g.timings()
    res = g.get_accelerations()
    if components:
        res.update(g.get_components(tree=True, pm=True))
    g.close()
    return dict(total_s=total, step_ms=step_ms, timings=tim), res


def host_io_run(A, pos, mass, eps, old_acc, steps, pinned=False):
    """The drop-in as AREPO's glue drives it every step: positions, masses and softening types from host
    arrays (agrav_set_particles), OldAcc from the host (agrav_set_old_acc), the step, and the
    accelerations and OldAcc back to host arrays (agrav_get_accelerations).  Host memory: pageable, as
    the glue's myrealloc buffers, or (pinned) staging arrays from agrav_host_alloc that the glue fills
    instead.  Times each part per step."""
    g = A.Gravity(make_params(A, A.FP32, A.WALK_GROUP, eps))
    n = len(pos)
    if pinned:
        hp, hm, st = A.host_empty((n, 3)), A.host_empty(n), A.host_empty(n, np.uint8)
        hp[:], hm[:], st[:] = pos, mass, 0
        pos, mass = hp, hm
        oa = A.host_empty(n, np.float32)
        oa[:] = old_acc
        out = {"acc": A.host_empty((n, 3)), "old_acc": A.host_empty(n, np.float32)}
    else:
        st = np.zeros(n, np.uint8)
        oa = old_acc.copy()
        out = {"acc": np.empty((n, 3)), "old_acc": np.empty(n, np.float32)}
    t = {"h2d": [], "device": [], "d2h": [], "total": []}
    for k in range(steps + 1):
        t0 = time.perf_counter()
        g.set_particles(pos, mass, softtype=st)
        g.set_old_acc(oa)
        t1 = time.perf_counter()
        g.compute_grav_accelerations(first_step=False)
        t2 = time.perf_counter()
        r = g.get_accelerations(fields=("acc", "old_acc"), out=out)
        t3 = time.perf_counter()
        oa = r["old_acc"]
        if k > 0:  # the first call is the warm-up
            for key, v in (("h2d", t1 - t0), ("device", t2 - t1), ("d2h", t3 - t2), ("total", t3 - t0)):
                t[key].append(1e3 * v)
    g.close()
    med = {k: float(np.median(v)) for k, v in t.items()}
    return {"steps": steps, "ms_per_step": med["total"], "particles_per_s": len(pos) / (med["total"] * 1e-3),
            "h2d_ms": med["h2d"], "device_ms": med["device"], "d2h_ms": med["d2h"],
            "bytes_in": int(len(pos) * (24 + 8 + 1 + 4)), "bytes_out": int(len(pos) * (24 + 4)),
            "host_memory": "page-locked (agrav_host_alloc)" if pinned else "pageable numpy arrays"}


def substep_run(A, pos, mass, eps, old_acc, fractions, steps):
    """Hierarchical sub-steps on the same particle set: for each active fraction, a seeded random active
    list, then `steps` timed compute_grav_accelerations calls (FP32 group walk) with (a) the partial tree of
    HIERARCHICAL_GRAVITY (builds insert only the active particles, OldAcc kept) and (b) the tree of every
    particle with the active list as walk targets.  (b) runs the PM of every particle on each call; (a) keeps
    the last one while the particles are unchanged, as GravPM between PM steps."""
    out = {}
    n = len(pos)
    for f in fractions:
        act = np.sort(np.random.Generator(np.random.Philox(97)).choice(n, int(n * f), replace=False)).astype(np.int32)
        row = {"active": int(len(act))}
        for mode in ("partial_tree", "full_tree_active_targets"):
            g = A.Gravity(make_params(A, A.FP32, A.WALK_GROUP, eps, group_size=0))  # group size by target density
            g.set_particles(pos, mass)
            g.set_old_acc(old_acc)
            if mode == "partial_tree":
                g.set_partial_tree(True, set_oldacc=False)
            g.set_active(act)
            g.compute_grav_accelerations(first_step=False)  # warmup
            t0 = time.perf_counter()
            for _ in range(steps):
                g.compute_grav_accelerations(first_step=False)
            dt = (time.perf_counter() - t0) / steps
            tim = g.timings()
            g.close()
            row[mode] = {"ms_per_step": 1e3 * dt, "treebuild_ms": tim["treebuild"], "treewalk_ms": tim["treewalk"],
                         "pm_ms": tim["pm_grav"], "interactions_per_active": tim["interactions"] / max(len(act), 1)}
        out["%g" % f] = row
    return out


def serving_run(A, precision, walk, slots, rate, seconds, seed, max_requests=None):
    """Poisson stream of isolated Hernquist solves (N log-normal, median 5e4, sigma 1.2 in ln, clipped to
    [256, 4M]) submitted in real time to a continuous-batching server; returns latency percentiles and
    throughput.  Particle sets are random subsets of one pre-drawn 4M-particle Hernquist sample (a subset
    of independent draws is itself a Hernquist sample)."""
    import threading
    from arepo_amd import ics
    g = ics.rng(seed)
    pool_n = 1 << 22
    pool = ics._isotropic(ics.hernquist_radii(pool_n, 1.0, g), g)
    t_arr = np.cumsum(g.exponential(1.0 / rate, int(rate * seconds * 1.5) + 10))
    t_arr = t_arr[t_arr < seconds]
    if max_requests is not None:
        t_arr = t_arr[:max_requests]
    sizes = np.clip(np.exp(np.log(5e4) + 1.2 * g.standard_normal(len(t_arr))), 256, 4e6).astype(np.int64)
    starts = g.integers(0, pool_n - sizes + 1)
    eps = (1.0 + np.sqrt(2.0)) / 50.0  # 1/50 of the Hernquist half-mass radius (a = 1)
    p = A.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), precision=precision, walk=walk,
                      group_size=0, leaf_bucket=4, eval_potential=True)
    lat, runs = [], []
    with A.Server(p, slots=slots, hbm_budget_bytes=64 << 30) as srv:
        # warm the slots up (device buffers, first-launch costs)
        for t in [srv.submit(pool[:20000], np.full(20000, 1 / 20000)) for _ in range(slots)]:
            srv.wait(t)
        tickets = []
        done = threading.Event()
        results = []

        def collector():
            k = 0
            while not done.is_set() or k < len(tickets):
                if k < len(tickets):
                    r = srv.wait(tickets[k])
                    results.append(r)
                    k += 1
                else:
                    time.sleep(0.0005)

        th = threading.Thread(target=collector)
        th.start()
        t0 = time.perf_counter()
        for ta, n, s0 in zip(t_arr, sizes, starts):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            tickets.append(srv.submit(pool[s0:s0 + n], np.full(n, 1.0 / n)))
        done.set()
        th.join()
        t_end = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([r["total_ms"] for r in results])
    runs = np.array([r["run_ms"] for r in results])
    return {"requests": int(len(results)), "arrival_rate_per_s": rate, "window_s": float(seconds),
            "makespan_s": t_end, "solves_per_s": len(results) / t_end,
            "particles_per_s": float(sizes[:len(results)].sum() / t_end),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "run_ms": {"p50": float(np.percentile(runs, 50)), "p99": float(np.percentile(runs, 99))},
            "mean_n": float(sizes[:len(results)].mean()), "slots": slots, "peak_running": st["peak_running"],
            "failed": st["failed"]}


def batch_run(A, nsets, steps, seed, cpu_particles, threads, err_sample):
    """Batched potential + acceleration solves of independent isolated groups (unbinding): particle counts
    log-uniform on [32, 200000], NFW- or Hernquist-like, softening = half-mass radius / 50, FP32 offsets
    from each group's centre; Barnes-Hut with theta = 0.7 as the reference's unbinding walk.  One tree
    build and one persistent walk launch per solve, group size and leaf bucket autotuned per size class.
    Baselines: the same sets solved one launch sequence per set (sample), and the oracle's restatement
    of the reference unbinding walk on the host (sample)."""
    from concurrent.futures import ThreadPoolExecutor
    from arepo_amd import ics
    log("batch: generating %d groups" % nsets)
    off, rel, mass, eps, cen = ics.unbinding_batch(nsets, seed=seed)
    n = int(off[-1])
    sizes = np.diff(off)
    p = A.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,), precision=A.FP32, walk=A.WALK_GROUP,
                      err_tol_theta=0.7, opening=0, leaf_bucket=4, eval_potential=True)
    res = {"n_sets": nsets, "n_particles": n, "n_min": int(sizes.min()), "n_max": int(sizes.max()),
           "criterion": "barnes-hut theta 0.7 (subfind_loctree.c:639)", "precision": "fp32"}
    g = A.Gravity(p)
    t0 = time.perf_counter()
    g.batch_set_particles(off, rel, mass, eps, centres=cen)
    res["upload_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    tun = g.batch_autotune()
    res["autotune_s"] = time.perf_counter() - t0
    res["tuning"] = {k: tun[k] for k in ("group_size", "leaf_bucket", "class_max_n", "sample_particles", "best_ms")}
    log("batch: autotune %.1f s -> %s" % (res["autotune_s"], json.dumps(res["tuning"])))
    g.compute_grav_accelerations(first_step=True)  # warm-up
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        g.compute_grav_accelerations(first_step=True)
        step_ms.append(1e3 * (time.perf_counter() - ts))
    total = time.perf_counter() - t0
    tim = g.timings()
    res.update({"steps": steps, "ms_per_solve": 1e3 * total / steps, "particles_per_s": n * steps / total,
                "interactions_per_s": tim["interactions"] * steps / total,
                "interactions_per_particle": tim["interactions"] / n,
                "stage_ms": {"treebuild": tim["treebuild"], "treewalk": tim["treewalk"], "finalize": tim["finalize"]},
                "walk_kernel_ms": tim["walk_kernel"], "numnodes": tim["numnodes"],
                "walk_fallback_groups": tim["walk_retries"]})
    log("batch: %.1f ms/solve (%.3g particles/s), build %.1f ms, walk %.1f ms"
        % (res["ms_per_solve"], res["particles_per_s"], tim["treebuild"], tim["treewalk"]))
    out = g.get_accelerations()
    # errors against FP64 direct summation over each target's own group
    sel = np.sort(np.random.Generator(np.random.Philox(97)).choice(n, min(err_sample, n), replace=False))
    sel = sel.astype(np.int32)
    ft = g.gravity_forcetest(sel)
    ea = rel_err(out["acc"][sel], ft["acc"])
    ep = np.abs(out["pot"][sel] - ft["pot"]) / np.abs(ft["pot"])
    res["error_vs_direct"] = {"acc": {"median": float(np.median(ea)), "p99": float(np.percentile(ea, 99))},
                              "pot": {"median": float(np.median(ep)), "p99": float(np.percentile(ep, 99))},
                              "targets": int(len(sel))}
    g.close()
    # FP64 group walk on the same batch
    p64 = A.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,), precision=A.FP64,
                        walk=A.WALK_GROUP, err_tol_theta=0.7, opening=0, leaf_bucket=4, eval_potential=True)
    with A.Gravity(p64) as g64:
        g64.batch_set_particles(off, rel, mass, eps, centres=cen)
        g64.batch_set_tuning(tun["group_size"], tun["leaf_bucket"])
        g64.compute_grav_accelerations(first_step=True)
        t0 = time.perf_counter()
        g64.compute_grav_accelerations(first_step=True)
        res["fp64_group_ms_per_solve"] = 1e3 * (time.perf_counter() - t0)
    # one launch sequence per set (the per-group loop of the reference, on the GPU), every 16th set
    samp = np.arange(0, nsets, 16)
    with A.Gravity(p) as g1:
        g1.set_particles(rel[off[0]:off[1]].astype(np.float64) + cen[0], mass[off[0]:off[1]])
        g1.compute_grav_accelerations()
        t_sets = 0.0
        for s in samp:
            pos = rel[off[s]:off[s + 1]].astype(np.float64) + cen[s]
            g1.set_softenings([eps[s]])
            g1.set_particles(pos, mass[off[s]:off[s + 1]])
            t0 = time.perf_counter()
            g1.compute_grav_accelerations(first_step=True)
            t_sets += time.perf_counter() - t0
    n_samp = int(sizes[samp].sum())
    res["per_set_launches"] = {"sets": int(len(samp)), "particles": n_samp, "seconds": t_sets,
                               "particles_per_s": n_samp / t_sets,
                               "batched_speedup": res["particles_per_s"] / (n_samp / t_sets)}
    # the oracle's unbinding walk (FP64 restatement of subfind_loctree.c:564-730) on the host, by group,
    # groups in parallel: every k-th group until about cpu_particles particles
    if cpu_particles > 0:
        import oracle as O
        O.lib()
        stride = max(1, int(np.ceil(n / cpu_particles)))
        cs = np.arange(0, nsets, stride)

        def one(s):
            pos = rel[off[s]:off[s + 1]].astype(np.float64) + cen[s]
            return O.group_potential_bh(pos, mass[off[s]:off[s + 1]], 2.8 * eps[s], 0.7)[0]

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            pots = list(ex.map(one, cs))
        t_cpu = time.perf_counter() - t0
        n_cpu = int(sizes[cs].sum())
        res["cpu_baseline"] = {"value": n_cpu / t_cpu, "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": "oracle unbinding walk (tree build + Barnes-Hut potential) of every "
                                         "%d-th group: %d groups, %d particles, %.1f s" % (stride, len(cs), n_cpu, t_cpu)}
        res["speedup_vs_cpu"] = res["particles_per_s"] / res["cpu_baseline"]["value"]
        # the oracle walk's own potential error on the sampled targets that fall in its groups
        in_cs = np.isin(np.searchsorted(off, sel, side="right") - 1, cs)
        if in_cs.any():
            lookup = {int(s): q for q, s in enumerate(cs)}
            gsel = np.searchsorted(off, sel[in_cs], side="right") - 1
            pc = np.array([pots[lookup[int(s)]][i - off[s]] for s, i in zip(gsel, sel[in_cs])])
            e = np.abs(pc - ft["pot"][in_cs]) / np.abs(ft["pot"][in_cs])
            res["error_vs_direct"]["cpu_oracle_pot"] = {"median": float(np.median(e)),
                                                        "p99": float(np.percentile(e, 99)), "targets": int(len(e))}
            eg = ep[in_cs]
            res["error_vs_direct"]["gpu_pot_same_targets"] = {"median": float(np.median(eg)),
                                                              "p99": float(np.percentile(eg, 99))}
    return res


def unbind_run(A, ngroups, seed, cpu_particles, threads):
    """SUBFIND unbinding (subfind_unbind, subfind_serial.c:429-637) of candidate groups: the batch
    workload's groups (log-uniform 32 .. 200000 members, NFW- or Hernquist-like) with isotropic velocities
    at U(0, 1.3) x escape speed, 5 % interlopers at 2-4 x escape speed and a bulk velocity per group.
    All groups advance one pass per round: one batched FP32 potential solve (autotuned group walk) for
    the groups in a full pass, one workgroup per group for the rest.  Baseline: the oracle's restatement
    of subfind_unbind on the host (groups in parallel, sample), and the bound sets of that sample compared
    with the GPU's."""
    from concurrent.futures import ThreadPoolExecutor
    from arepo_amd import ics
    log("unbind: generating %d candidate groups" % ngroups)
    off, pos, vel, mass, eps = ics.unbinding_candidates(ngroups, seed=seed)
    n = int(off[-1])
    sizes = np.diff(off)
    p = A.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,), precision=A.FP32, walk=A.WALK_GROUP,
                      err_tol_theta=0.7, opening=0, leaf_bucket=4, eval_potential=True)
    res = {"n_groups": ngroups, "n_particles": n, "n_min": int(sizes.min()), "n_max": int(sizes.max()),
           "criterion": "barnes-hut theta 0.7 (ErrTolThetaSubfind, subfind_loctree.c:639)", "precision": "fp32",
           "des_link_ngb": 20}
    with A.Gravity(p) as g:
        # the batch autotuner on the initial member sets sets the walk's group size and leaf bucket
        g.batch_set_particles(off, pos.astype(np.float32), mass, eps)
        tun = g.batch_autotune()
        res["tuning"] = {k: tun[k] for k in ("group_size", "leaf_bucket")}
        g.subfind_unbind(off, pos, vel, mass, eps, des_link_ngb=20)  # warm-up (allocations)
        t0 = time.perf_counter()
        r = g.subfind_unbind(off, pos, vel, mass, eps, des_link_ngb=20)
        wall = time.perf_counter() - t0
        tim = g.timings()
    left = int(r["len"].sum())
    res.update({"seconds": wall, "particles_per_s": n / wall, "members_left": left,
                "bound_fraction": left / n, "rounds": tim["unbind_rounds"], "potential_solves": tim["unbind_solves"],
                "solve_ms": tim["unbind_solve_ms"], "pass_ms": tim["unbind_pass_ms"],
                "passes_per_group": {"median": float(np.median(r["iterations"])),
                                     "max": int(r["iterations"].max())}})
    log("unbind: %.3f s (%.3g particles/s), %d rounds, %d potential solves (%.0f ms), passes %.0f ms, %.3f bound"
        % (wall, n / wall, tim["unbind_rounds"], tim["unbind_solves"], tim["unbind_solve_ms"],
           tim["unbind_pass_ms"], left / n))
    if cpu_particles > 0:
        import oracle as O
        O.lib()
        stride = max(1, int(np.ceil(n / cpu_particles)))
        cs = np.arange(0, ngroups, stride)
        up = O.make_unbind_params(theta=0.7, des_link_ngb=20)

        def one(s):
            a, b = int(off[s]), int(off[s + 1])
            return O.subfind_unbind(up, pos[a:b], vel[a:b], mass[a:b], eps[s])

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            outs = list(ex.map(one, cs))
        t_cpu = time.perf_counter() - t0
        n_cpu = int(sizes[cs].sum())
        res["cpu_baseline"] = {"value": n_cpu / t_cpu, "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": "oracle subfind_unbind of every %d-th group: %d groups, %d particles, %.1f s"
                                         % (stride, len(cs), n_cpu, t_cpu)}
        res["speedup_vs_cpu"] = res["particles_per_s"] / res["cpu_baseline"]["value"]
        ob = np.concatenate([o["bound"] for o in outs])
        gb = np.concatenate([r["bound"][off[s]:off[s + 1]] for s in cs])
        res["bound_set_vs_oracle"] = {"members": int(len(ob)), "differ": int((ob != gb).sum()),
                                      "differ_fraction": float((ob != gb).mean()),
                                      "groups_same_len": int(sum(int(o["len"] == r["len"][s])
                                                                 for o, s in zip(outs, cs))),
                                      "groups": int(len(cs))}
        log("unbind: cpu %.3g particles/s, bound sets differ for %d of %d members"
            % (res["cpu_baseline"]["value"], res["bound_set_vs_oracle"]["differ"], len(ob)))
    return res


PMC_PROFILE = os.path.join(ROOT, "profiles", "k_walk_group_treepm128_pmc.md")


def profiled_counters():
    """From the committed rocprofv3 counter summary of the walk kernel: HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE, each scaled by a copy of known size) and the VALU issue time as a
    fraction of the kernel time; None where absent."""
    got = {}
    try:
        for line in open(PMC_PROFILE):
            for key in ("traffic_bytes_per_launch", "valu_issue_fraction", "l2_hit_rate", "kernel_ms"):
                if line.startswith(key + ":"):
                    got[key] = float(line.split(":")[1])
    except OSError:
        pass
    return got


def copy_probe(nbytes=256 << 20):
    """One device-to-device copy of known size (the runtime's copy kernel), used to scale the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy probe: %d bytes device-to-device" % nbytes)


def cpu_baseline(pos, mass, eps, old_acc, sel, threads):
    """Oracle (FP64 AREPO restatement) on the host: full tree build and PM, walk timed on a random
    sample of targets and scaled to all particles."""
    import oracle as O
    p = O.make_params(periodic=True, box_size=1.0, pmgrid=128, asmth=1.25, rcut=4.5, err_tol_theta=0.7,
                      err_tol_force_acc=0.0025, opening=1, soft=(eps,))
    n = len(pos)
    t0 = time.perf_counter()
    acc_pm, _ = O.pm_periodic(p, pos, mass)
    t1 = time.perf_counter()
    tree = O.Tree(p, pos, mass)
    t2 = time.perf_counter()
    sample = len(sel)
    acc, pot, nint, tot = tree.walk(False, old_acc=old_acc, targets=sel, nthreads=threads)
    t3 = time.perf_counter()
    t_walk_full = (t3 - t2) * n / sample
    t_total = (t1 - t0) + (t2 - t1) + t_walk_full
    # one core: the PM and the build are serial already; the walk of every 16th sample target on one thread
    sub = sel[::16]
    t4 = time.perf_counter()
    tree.walk(False, old_acc=old_acc, targets=sub, nthreads=1)
    t5 = time.perf_counter()
    t_total_1 = (t1 - t0) + (t2 - t1) + (t5 - t4) * n / len(sub)
    return dict(value=n / t_total, t_pm=t1 - t0, t_build=t2 - t1, t_walk_sample=t3 - t2, sample=sample,
                ia_per_part=tot / sample, sel=sel, acc_tree=acc, pot_tree=pot, acc_pm=acc_pm, params=p,
                value_1core=n / t_total_1, sample_1core=len(sub), t_walk_sample_1core=t5 - t4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128 ** 3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--err-sample", type=int, default=16384)
    ap.add_argument("--cpu-sample", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--substep-steps", type=int, default=5, help="timed calls per active fraction (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the Poisson request stream (the serving workload: 60 s)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--batch-sets", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-cpu-particles", type=int, default=8 << 20)
    ap.add_argument("--batch-only", action="store_true", help="run only the batched-unbinding leg (profiling aid)")
    ap.add_argument("--unbind-groups", type=int, default=4096)
    ap.add_argument("--unbind-cpu-particles", type=int, default=2 << 20)
    ap.add_argument("--unbind-only", action="store_true", help="run only the unbinding-loop leg (profiling aid)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the gravity step of one particle set runs on a single MI355X "
              "(the path is not sharded across GPUs here); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from arepo_amd import agrav as A
    from arepo_amd import ics

    if args.batch_only:
        res = batch_run(A, args.batch_sets, args.batch_steps, 4, 0 if args.no_cpu else args.batch_cpu_particles,
                        min(os.cpu_count() or 1, 16), args.err_sample)
        print(json.dumps({"batched_unbinding": res}), flush=True)
        return

    if args.unbind_only:
        res = unbind_run(A, args.unbind_groups, 6, 0 if args.no_cpu else args.unbind_cpu_particles,
                         min(os.cpu_count() or 1, 16))
        print(json.dumps({"subfind_unbinding": res}), flush=True)
        return

    log("generating clustered 128^3 set (seed %d)" % args.seed)
    pos, mass, eps = ics.clustered_box(args.n, seed=args.seed)
    n = len(pos)
    # one seeded target subset serves the error measurement and the CPU sample
    sel = np.sort(np.random.Generator(np.random.Philox(99)).choice(n, min(args.err_sample, n), replace=False))
    sel = sel.astype(np.int32)

    if args.copy_probe:
        copy_probe()

    # FP32 group walk: the measured path
    log("FP32 group walk: warmup %d, steps %d" % (args.warmup, args.steps))
    r32, res32 = run_path(A, pos, mass, eps, A.FP32, A.WALK_GROUP, args.steps, args.warmup, components=True)
    tim = r32["timings"]
    ms_per_step = 1e3 * r32["total_s"] / args.steps
    value = n * args.steps / r32["total_s"]
    inter = tim["interactions"]
    log("fp32 group: %.2f ms/step, walk %.2f ms, build %.2f ms, pm %.2f ms, ia/part %.1f"
        % (ms_per_step, tim["treewalk"], tim["treebuild"], tim["pm_grav"], inter / n))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded clustered particle set generated in-process)",
        "config": {"workload": "treepm_128cubed_clustered", "n_particles": n, "pmgrid": 128, "asmth": 1.25,
                   "rcut": 4.5, "err_tol_force_acc": 0.0025, "opening": "relative", "softening": eps,
                   "precision": "fp32", "walk": "group64", "leaf_bucket": 4,
                   "evalpotential": False, "seed": args.seed},
        "interactions_per_s": inter * args.steps / r32["total_s"],
        "interactions_per_particle": inter / n,
        "latency_ms": {"p50": float(np.percentile(r32["step_ms"], 50)),
                       "p99": float(np.percentile(r32["step_ms"], 99))},
        "stage_ms": {"treebuild": tim["treebuild"], "treewalk": tim["treewalk"], "pm_grav": tim["pm_grav"],
                     "finalize": tim["finalize"]},
        "numnodes": tim["numnodes"],
        "walk_fallback_groups": tim["walk_retries"],
        # group-walk work per step: stack entries evaluated, and tile entries each streamed by all 64 lanes
        # (interactions / (64 x tile entries) is the fraction of tile lane work that is an interaction)
        "walk_popped": tim["walk_popped"],
        "tile_entries": tim["tile_entries"],
    }

    # roofline of the dominant kernel (the walk): algorithmic bytes = one 16-byte source record per
    # interaction + 52 bytes per target (position, mass, softening/OldAcc in, acc + pot out)
    walk_s = tim["walk_kernel"] * 1e-3
    alg_bytes = inter * 16 + n * 52
    alg_rate = alg_bytes / walk_s if walk_s > 0 else 0.0
    prof = profiled_counters()
    traffic = prof.get("traffic_bytes_per_launch")
    # the counters come from a committed profile of this kernel at this workload: they stand for this run
    # only while its kernel time agrees with the profile's
    prof_ms = prof.get("kernel_ms")
    stale = prof_ms is not None and abs(tim["walk_kernel"] - prof_ms) > 0.1 * prof_ms
    if stale:
        log("WARNING: walk kernel %.3f ms here, %.3f ms in %s: the profiled traffic may be stale"
            % (tim["walk_kernel"], prof_ms, os.path.relpath(PMC_PROFILE, ROOT)))
    # Node and particle records are reused from L2 and from the LDS tile by all lanes of a group, so
    # the algorithmic rate can pass the HBM peak.  Then the rate of the measured HBM traffic (rocprofv3
    # FETCH_SIZE + WRITE_SIZE of this kernel at this workload) is the roofline figure, flagged
    # cache-amplified, and the algorithmic rate is reported beside it; a fraction above 1 is never
    # claimed.  The kernel's real limit is VALU issue, reported from the same profile.
    amplified = alg_rate > HBM_PEAK and traffic is not None and walk_s > 0
    achieved = traffic / walk_s if amplified else alg_rate
    flops = inter * 28  # TreePM short-range monopole, 28 flops per interaction (sqrt and divide as 1)
    out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK, "traffic": traffic, "cache_amplified": bool(amplified),
                       "kernel": "k_walk_group<float,PM>", "kernel_ms": tim["walk_kernel"],
                       "algorithmic": {"bytes_per_launch": alg_bytes, "rate_GBps": alg_rate / 1e9,
                                       "frac_of_peak": alg_rate / HBM_PEAK},
                       "secondary": {"fp32_tflops": flops / walk_s / 1e12 if walk_s > 0 else 0.0,
                                     "fp32_peak_tflops": FP32_PEAK / 1e12,
                                     "fp32_frac": flops / walk_s / FP32_PEAK if walk_s > 0 else 0.0,
                                     "valu_issue_fraction_of_kernel_time": prof.get("valu_issue_fraction"),
                                     "l2_hit_rate": prof.get("l2_hit_rate")},
                       "profiled": os.path.relpath(PMC_PROFILE, ROOT) if prof else None,
                       "profiled_kernel_ms": prof_ms, "profile_stale": bool(stale)}

    # whole-step algorithmic bytes, the per-solve formula of the path's scope table:
    # I*16 + N*52 + tree (keys+indices, 4 passes read+write; 48 B per node read+write)
    # + PM (deposit and readout per particle, 24 passes of 4 B over the G^3 mesh)
    G = 128
    b_tree = n * 12 * 2 * 4 + tim["numnodes"] * 48 * 2
    b_pm = n * 32 + n * 40 + G ** 3 * 4 * 24
    b_step = inter * 16 + n * 52 + b_tree + b_pm
    out["roofline_step"] = {"bytes_per_step": b_step, "rate_GBps": b_step / (ms_per_step * 1e-3) / 1e9,
                            "frac_of_hbm_peak": b_step / (ms_per_step * 1e-3) / HBM_PEAK,
                            "cache_amplified": True,
                            "note": "I*16 B dominates; node and particle records are reused from L2 and LDS, so this "
                                    "is not a measured HBM rate (roofline.traffic is)"}

    # the same path with EVALPOTENTIAL (potentials summed too)
    rpot, _ = run_path(A, pos, mass, eps, A.FP32, A.WALK_GROUP, 3, 0, old_acc=res32["old_acc"], pot=True)
    out["fp32_group_with_potential"] = {"ms_per_step": 1e3 * rpot["total_s"] / 3,
                                        "walk_ms": rpot["timings"]["treewalk"]}

    # AREPO's own calling pattern: its domain decomposition reorders P[] by Peano-Hilbert key
    # (domain.c:140, peano_hilbert_order), so the gravity step receives its particles in about the
    # tree's order, and the build's gather and the walk's caller-order writes become near-sequential.
    # The same step on the same particles stored in the tree's order (`value` stays on the
    # generator's order, in which the gather reads at random).
    gh = A.Gravity(make_params(A, A.FP32, A.WALK_GROUP, eps))
    gh.set_particles(pos, mass)
    gh.construct_forcetree()
    perm, _ = gh.get_order()
    gh.close()
    rh, _ = run_path(A, pos[perm], mass[perm], eps, A.FP32, A.WALK_GROUP, args.steps, 1,
                     old_acc=res32["old_acc"][perm])
    th = rh["timings"]
    out["fp32_group_tree_ordered_input"] = {
        "ms_per_step": 1e3 * rh["total_s"] / args.steps, "particles_per_s": n * args.steps / rh["total_s"],
        "stage_ms": {"treebuild": th["treebuild"], "treewalk": th["treewalk"], "pm_grav": th["pm_grav"]},
        "input_order": "the library's Peano-Hilbert order of the particles, as P[] after AREPO's domain decomposition"}
    log("fp32 group, caller arrays in the tree's order: %.2f ms/step, build %.2f ms"
        % (out["fp32_group_tree_ordered_input"]["ms_per_step"], th["treebuild"]))

    if args.substep_steps > 0:
        log("hierarchical sub-steps: partial tree vs full tree with active targets")
        out["hierarchical_substeps"] = substep_run(A, pos, mass, eps, res32["old_acc"], (0.5, 0.125, 0.03125),
                                                   args.substep_steps)
        log("sub-steps: %s" % json.dumps(out["hierarchical_substeps"]))

    res64p = None
    if not args.no_fp64:
        log("FP64 plain walk (one lane per target)")
        r64p, res64p = run_path(A, pos, mass, eps, A.FP64, A.WALK_PLAIN, 1, 0, old_acc=res32["old_acc"],
                                components=True)
        log("FP64 group walk")
        r64g, res64g = run_path(A, pos, mass, eps, A.FP64, A.WALK_GROUP, 2, 0, old_acc=res32["old_acc"])
        out["fp64_plain"] = {"ms_per_step": 1e3 * r64p["total_s"], "walk_ms": r64p["timings"]["treewalk"],
                             "interactions": r64p["timings"]["interactions"]}
        out["fp64_group"] = {"ms_per_step": 1e3 * r64g["total_s"] / 2, "walk_ms": r64g["timings"]["treewalk"],
                             "interactions": r64g["timings"]["interactions"]}
        out["speedup_fp32_vs_fp64_plain"] = out["fp64_plain"]["ms_per_step"] / ms_per_step
        # the interactions the reference's per-particle walk makes (the FP64 plain count), per second of
        # the FP32 step: the group walk evaluates more pairs, each far cheaper
        out["reference_equivalent_interactions_per_s"] = r64p["timings"]["interactions"] / (ms_per_step * 1e-3)

    log("drop-in with host arrays (set_particles + set_old_acc + step + get_accelerations)")
    out["fp32_group_host_io"] = host_io_run(A, pos, mass, eps, res32["old_acc"], max(3, min(args.steps, 10)))
    out["fp32_group_host_io_pinned"] = host_io_run(A, pos, mass, eps, res32["old_acc"], max(3, min(args.steps, 10)),
                                                   pinned=True)
    log("host io: %s" % json.dumps(out["fp32_group_host_io"]))

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        extra = np.random.Generator(np.random.Philox(98)).choice(n, min(args.cpu_sample, n), replace=False)
        cpu_sel = np.union1d(sel, extra).astype(np.int32)
        log("CPU oracle baseline: %d threads, walk sample %d" % (threads, len(cpu_sel)))
        cpu = cpu_baseline(pos, mass, eps, res32["old_acc"], cpu_sel, threads)
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": "full tree build + 128^3 PM on the host (the oracle's own complex radix-2 "
                                         "FFT, no FFTW: slower than the reference's PM leg), relative walk of %d random "
                                         "targets scaled to N (build %.2f s, pm %.2f s, walk sample %.2f s)"
                                         % (cpu["sample"], cpu["t_build"], cpu["t_pm"], cpu["t_walk_sample"])}
        out["speedup_vs_cpu"] = value / cpu["value"]
        out["cpu_baseline_1core"] = {"value": cpu["value_1core"], "unit": "particles/s", "cores": 1, "kind": "port",
                                     "sample": "the same build and PM; one-thread walk of %d of the sample targets "
                                               "(%.2f s) scaled to N" % (cpu["sample_1core"],
                                                                        cpu["t_walk_sample_1core"])}

    if not args.no_errors:
        log("FP64 direct summation (Ewald) on %d targets" % len(sel))
        g = A.Gravity(make_params(A, A.FP64, A.WALK_GROUP, eps))
        g.set_particles(pos, mass)
        t0 = time.perf_counter()
        ft = g.gravity_forcetest(sel)
        log("direct sum took %.1f s" % (time.perf_counter() - t0))
        g.close()
        errs = {}
        a_norm = np.linalg.norm(ft["acc"], axis=1)

        def stat(e):
            return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99))}

        def split(acc_tree, acc_pm):
            # the tree part against FORCETEST's exact short-range sum, the PM part against the long-range
            # rest (gravtree_forcetest.c:629-638), both relative to |a|
            return {"tree": stat(np.linalg.norm(acc_tree - ft["acc_short"], axis=1) / a_norm),
                    "pm": stat(np.linalg.norm(acc_pm - (ft["acc"] - ft["acc_short"]), axis=1) / a_norm)}

        for name, res in (("fp32_group", res32), ("fp64_plain", res64p)):
            if res is None:
                continue
            errs[name] = stat(rel_err(res["acc"][sel], ft["acc"]))
            if res.get("acc_tree") is not None:
                errs[name]["split"] = split(res["acc_tree"][sel], res["acc_pm"][sel])
        if cpu is not None:
            # the oracle's own TreePM error on its sample targets (tree part from the oracle, PM from the oracle)
            pos_in = np.searchsorted(cpu["sel"], sel)
            errs["cpu_oracle"] = stat(rel_err(cpu["acc_tree"][pos_in] + cpu["acc_pm"][sel], ft["acc"]))
            errs["cpu_oracle"]["split"] = split(cpu["acc_tree"][pos_in], cpu["acc_pm"][sel])
        out["force_error_vs_direct"] = errs
        out["force_error_sample"] = int(args.err_sample)

    if args.serve_seconds > 0:
        log("serving: Poisson %.0f req/s for %.0f s, FP32 group walk, 4 slots" % (args.serve_rate, args.serve_seconds))
        out["serving"] = serving_run(A, A.FP32, A.WALK_GROUP, 4, args.serve_rate, args.serve_seconds, seed=5)
        log("serving: %s" % json.dumps(out["serving"]))
        # baseline: the plain one-solve-per-launch FP64 path (one slot), on the first 2 s of the same stream
        base_s = min(2.0, args.serve_seconds)
        log("serving baseline: FP64 plain walk, 1 slot, first %.0f s of the stream" % base_s)
        out["serving_fp64_plain_1slot"] = serving_run(A, A.FP64, A.WALK_PLAIN, 1, args.serve_rate, base_s, seed=5)

    if args.batch_sets > 0:
        out["batched_unbinding"] = batch_run(A, args.batch_sets, args.batch_steps, 4,
                                             0 if args.no_cpu else args.batch_cpu_particles,
                                             min(os.cpu_count() or 1, 16), args.err_sample)
        log("batch: %s" % json.dumps(out["batched_unbinding"]))

    if args.unbind_groups > 0:
        out["subfind_unbinding"] = unbind_run(A, args.unbind_groups, 6,
                                              0 if args.no_cpu else args.unbind_cpu_particles,
                                              min(os.cpu_count() or 1, 16))
        log("unbind: %s" % json.dumps(out["subfind_unbinding"]))

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
