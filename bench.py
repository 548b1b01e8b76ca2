"""Benchmark: particle-steps/s of the Barnes-Hut step on one MI355X.

Workload (default): 500,000 particles, two exponential disks (scale length 1, height 0.1,
log-normal masses sigma 0.5) on a parabolic orbit 10 units apart, theta 0.7, eps^2 0.01,
63-bit keys (B = 21, 22 layers), dt 1e-3.  A step is one pass of the whole hot path over the
system: bounding cube + Morton keys, LSD radix sort, layered octree build (compact nodes and
the 128-B OctreeNode array), force traversal, integrator.

Prints ONE JSON line (rank 0).  The timed region runs K steps with HIP events around every
stage of every step on the library's stream; `value` is N*K / wall time of that region.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, packed / dual-issue
PEAK_FP32_TFLOPS_NONPACKED = 78.6
PEAK_HBM_GBS = 8000.0
FLOP_PER_INTERACTION = 20  # Nyland, Harris & Prins, GPU Gems 3 ch. 31 convention


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n-per-disk", type=int, default=250000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--cpu-steps", type=int, default=30,
                    help="oracle steps timed for the CPU baseline (about 10 s on 16 threads)")
    ap.add_argument("--cpu-steps-matched", type=int, default=12)
    ap.add_argument("--cpu-steps-1thread", type=int, default=2,
                    help="oracle steps timed on one host thread (about 10 s)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (<= 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="warmup + steps only, no JSON extras (for rocprofv3 runs)")
    ap.add_argument("--matched", action="store_true",
                    help="time the accuracy-matched configuration (quadrupoles at "
                         "--matched-theta)")
    ap.add_argument("--matched-theta", type=float, default=0.55,
                    help="opening angle of the accuracy-matched configuration: the largest on a "
                         "0.05 grid whose RMS force error on this workload stays <= 1e-3 "
                         "(0.55: 7.2e-4; 0.6: 1.02e-3)")
    ap.add_argument("--force-block", type=int, default=0, help="0: autotune")
    ap.add_argument("--top-cache", type=int, default=1,
                    help="1: the traversal reads the tree's top layers from LDS (blocks of "
                         "512/1024); 0: every node from device memory (for A/B profiles)")
    ap.add_argument("--xcd-chunk", type=int, default=-1, help="-1: autotune")
    ap.add_argument("--sort", choices=["seeded", "radix"], default="seeded",
                    help="the step's key sort: seeded by the previous step's order (default) or "
                         "a radix sort from scratch (for A/B profiles); the same order either way")
    ap.add_argument("--walk", choices=["lane", "wave"], default="lane",
                    help="the traversal: one walk per target (default) or one walk per wave over "
                         "the union of its targets' walks (for A/B profiles); the same forces")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--serve-rate", type=float, default=200.0,
                    help="serving stream: Poisson arrivals per second (200: the stream's own)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-steps-min", type=int, default=10)
    ap.add_argument("--serve-steps-max", type=int, default=1000)
    ap.add_argument("--serve-max-wait", type=float, default=5.0,
                    help="load shedding: refuse a request whose estimated wait exceeds this (s)")
    ap.add_argument("--serve-nominal-rate", type=float, default=2.0,
                    help="a second stream below capacity, for latency without queueing")
    ap.add_argument("--serve-nominal-seconds", type=float, default=15.0)
    ap.add_argument("--serve-drain-requests", type=int, default=48,
                    help="the first requests of the serving stream, all admitted, through both "
                         "forms (throughput on the same work, without shedding)")
    ap.add_argument("--serve-pool", type=int, default=32)
    ap.add_argument("--serve-capacity", type=int, default=8 << 20)
    ap.add_argument("--acc-steps", type=int, default=50,
                    help="timed steps of the accuracy-matched configuration")
    ap.add_argument("--acc-targets", type=int, default=4096)
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--large-n", type=int, default=4194304)
    ap.add_argument("--large-steps", type=int, default=20)
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="profile mode: also copy this many MiB device-to-device (a plain copy "
                         "of known size that scales the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


_T_START = time.perf_counter()


def progress(msg):
    """A progress line on stderr (the JSON result is the one line on stdout)."""
    print(f"[bench {time.perf_counter() - _T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def stage_bytes(n, V, passes=8, emit=True, quad=False, seeded=True):
    """Algorithmic HBM bytes per step of the stages.  seeded: the sort from the previous order
    (the keys computed in its place pass: the first stage only reads the positions for the
    cube)."""
    node = 66 + (128 if emit else 0) + (64 if quad else 0)  # write 33 + read 33 (+ OctreeNode, moments)
    return {
        # bbox read (+ Morton read, 2 key writes, iota for the radix sort)
        "bbox_morton": n * 16 if seeded else n * (16 + 16 + 8 + 8 + 4),
        # seeded: order 4 + position 16 read, slot 12 written; slot 12 read, key 8 + index 4 +
        # next order 4 written.  radix: per pass count reads 8, scatter moves 2 x 12
        "sort": n * 60 if seeded else n * 32 * passes,
        "build": n * 40 + V * node,                # leaf gather + per-node traffic
        "update": n * (16 + 16 + 16 + 4 + 16 + 16),
    }


def calibration_copy(nbytes):
    """One cold device-to-device copy of nbytes (HIP's blit kernel), for counter scaling."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    big = ctypes.c_void_p()  # evict the L2 by touching another buffer
    assert hip.hipMalloc(ctypes.byref(big), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(big, 2, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    for p in (src, dst, big):
        hip.hipFree(p)


def rms_rel(f, ref):
    d = np.asarray(f, np.float64)[:, :3] - ref
    return float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))


def bench_batched(astro, a):
    """256 independent Plummer systems, N log-uniform in [1024, 65536]: one batched step per
    graph replay vs the plain loop (each system in its own context, stepped in turn)."""
    sizes = astro.ic_batch_sizes(a.batch_systems, 1024, 65536, 7)
    parts = [astro.ic_plummer(int(n), 1000 + k) for k, n in enumerate(sizes)]
    # np.concatenate drops the structs' padding: back to the 32-B reference layouts
    pos = np.ascontiguousarray(np.concatenate([p[0] for p in parts]), astro.POSITION_MASS)
    vel = np.ascontiguousarray(np.concatenate([p[1] for p in parts]), astro.VELOCITY_RADIUS)
    col = np.ascontiguousarray(np.concatenate([p[2] for p in parts]), astro.COLOR_TYPE)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    ntot = int(off[-1])
    P = astro.params_physics(21)
    out = {"systems": len(sizes), "particles": ntot}
    with astro.Context(ntot, bits_per_axis=21, max_systems=len(sizes),
                       emit_octree_nodes=False) as ctx:
        ctx.set_params(P)
        ctx.upload_systems(pos, vel, col, off)
        out["autotuned_block"], out["autotuned_xcd_chunk"], _ = ctx.autotune(3)
        ctx.step(a.dt, 2)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(a.dt, a.batch_steps)
        ctx.sync()
        tb = time.perf_counter() - t0
        ctx.check()
    out["batched"] = {"value": ntot * a.batch_steps / tb, "ms_per_step": tb / a.batch_steps * 1e3,
                      "steps": a.batch_steps}
    ctxs = []
    for p_, v_, c_ in parts:
        c = astro.Context(len(p_), bits_per_axis=21, emit_octree_nodes=False)
        c.set_params(P)
        c.upload(p_, v_, c_)
        c.step(a.dt, 1)
        ctxs.append(c)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for _ in range(a.plain_steps):
        for c in ctxs:
            c.step(a.dt, 1)
    for c in ctxs:
        c.sync()
    tp = time.perf_counter() - t0
    for c in ctxs:
        c.check()
        c.close()
    out["plain_loop"] = {"value": ntot * a.plain_steps / tp, "ms_per_step": tp / a.plain_steps * 1e3,
                         "steps": a.plain_steps}
    out["batched_over_plain"] = out["batched"]["value"] / out["plain_loop"]["value"]
    # FP32 vs Kahan accumulation against the FP64 direct sum (on the GPU), on a few systems
    errs = {"fp32": [], "kahan": []}
    for k in np.linspace(0, len(parts) - 1, 4).astype(int):
        p_, v_, c_ = parts[k]
        tg = np.random.default_rng(k).choice(len(p_), min(2048, len(p_)), replace=False)
        for kahan in (0, 1):
            Pk = astro.params_physics(21)
            Pk.kahan = kahan
            with astro.Context(len(p_), bits_per_axis=21, emit_octree_nodes=False) as c:
                c.set_params(Pk)
                c.upload(p_, v_, c_)
                ref = c.direct_forces_fp64(tg)
                c.build_tree()
                c.compute_forces()
                errs["kahan" if kahan else "fp32"].append(rms_rel(c.forces()[tg], ref))
    out["rms_force_error_vs_fp64"] = {k: float(np.mean(v)) for k, v in errs.items()}
    out["rms_systems"] = "4 systems spread over the size range, 2048 random targets each"
    return out


def serving_pool(astro, a):
    rng = np.random.default_rng(5)
    pool_sizes = np.exp(rng.uniform(np.log(2048), np.log(1 << 20), a.serve_pool)).astype(int)
    pool = []
    for k, n in enumerate(pool_sizes):
        if k % 2 == 0:
            pool.append(astro.ic_plummer(int(n), 500 + k))
        else:
            pool.append(astro.ic_two_disks(max(1, int(n) // 2), 500 + k))
    return pool


def bench_serving(astro, a, pool, rate, seconds, seed, drain=0):
    """Poisson arrivals of Plummer / two-disk systems (N log-uniform in [2048, 2^20], steps
    uniform in [steps_min, steps_max]) for `seconds`, served with continuous batching; the
    same arrivals through the plain FIFO loop (one request at a time, each in a single-system
    context).  Both shed load the same way: a request whose estimated wait (the work ahead of
    it and its own, over the measured throughput) exceeds max_wait is refused.  Throughput is
    the completed work over the time from the first arrival to the last completion."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        reqs.append((len(reqs), t, int(rng.integers(0, len(pool))),
                     int(rng.integers(a.serve_steps_min, a.serve_steps_max + 1))))
        if drain and len(reqs) == drain:
            break
    max_wait = 0.0 if drain else a.serve_max_wait
    wk = lambda k, st: len(pool[k][0]) * st  # noqa: E731
    offered = sum(wk(k, st) for _, _, k, st in reqs)
    P = astro.params_physics(21)
    out = {"requests": len(reqs), "arrival_rate_per_s": rate, "arrival_seconds": seconds,
           "steps_range": [a.serve_steps_min, a.serve_steps_max],
           "max_wait_s": max_wait,
           "offered_particle_steps_per_s": offered / seconds,
           "pool": "%d seeded systems, N log-uniform in [2048, 1048576], Plummer and two-disk "
                   "alternating" % a.serve_pool}
    # served: the continuous-batching server
    srv = astro.Server(P, a.serve_capacity, 512, max_wait_s=max_wait)
    pending = list(reqs)
    lat, done_work, rejected = [], 0, 0
    t0 = time.perf_counter()
    t_note = t0
    progress(f"serving {rate:g}/s for {seconds:g} s: {len(reqs)} requests, served")
    while pending or srv.counters()["active"] or srv.counters()["queued"]:
        now = time.perf_counter() - t0
        if time.perf_counter() - t_note > 15:
            t_note = time.perf_counter()
            progress(f"  served: {len(reqs) - len(pending)} arrived, {len(lat)} completed, "
                     f"{rejected} refused")
        while pending and pending[0][1] <= now:
            rid, _, k, st = pending.pop(0)
            try:
                srv.submit(rid, pool[k][0], pool[k][1], pool[k][2], st)
            except astro.AstroError as e:
                if e.code != -7:  # ASTRO_ERR_BUSY
                    raise
                rejected += 1
        if srv.counters()["active"] or srv.counters()["queued"]:
            srv.tick(a.dt)
        else:
            time.sleep(0.0005)
        for rec in srv.poll():
            lat.append(rec["t_done"] - rec["t_submit"])
            done_work += rec["n"] * rec["steps_done"]
            srv.fetch(rec["id"], rec["n"])
    el = time.perf_counter() - t0 - reqs[0][1]
    srv.close()
    lat = np.array(lat)
    out["served"] = {"value": done_work / el, "seconds": el, "admitted": len(reqs) - rejected,
                     "rejected": rejected, "completed": len(lat),
                     "p50_latency_s": float(np.percentile(lat, 50)),
                     "p99_latency_s": float(np.percentile(lat, 99))}
    # plain loop: one request at a time, FIFO, each in a single-system context, the same
    # shedding rule with its own measured throughput
    ctx = astro.Context(int(max(len(p[0]) for p in pool)), bits_per_axis=21, emit_octree_nodes=False)
    ctx.set_params(P)
    pending = list(reqs)
    queue, lat, done_work, rejected, prate = [], [], 0, 0, 0.0
    t0 = time.perf_counter()
    t_note = t0
    progress("  plain loop")
    while pending or queue:
        now = time.perf_counter() - t0
        if time.perf_counter() - t_note > 15:
            t_note = time.perf_counter()
            progress(f"  plain: {len(reqs) - len(pending)} arrived, {len(lat)} completed, "
                     f"{rejected} refused")
        while pending and pending[0][1] <= now:
            r = pending.pop(0)
            ahead = sum(wk(q[2], q[3]) for q in queue) + wk(r[2], r[3])
            if max_wait > 0 and prate > 0 and ahead / prate > max_wait:
                rejected += 1
            else:
                queue.append(r)
        if not queue:
            if pending:
                time.sleep(max(0.0, pending[0][1] - (time.perf_counter() - t0)))
            continue
        rid, tarr, k, st = queue.pop(0)
        ts = time.perf_counter()
        ctx.upload(*pool[k])
        ctx.step(a.dt, st)
        ctx.sync()
        ctx.download()
        te = time.perf_counter()
        r_now = wk(k, st) / (te - ts)
        prate = r_now if prate == 0 else 0.7 * prate + 0.3 * r_now
        done_work += wk(k, st)
        lat.append(te - t0 - tarr)
    el = time.perf_counter() - t0 - reqs[0][1]
    ctx.check()
    ctx.close()
    lat = np.array(lat)
    out["plain_loop"] = {"value": done_work / el, "seconds": el, "admitted": len(reqs) - rejected,
                         "rejected": rejected, "completed": len(lat),
                         "p50_latency_s": float(np.percentile(lat, 50)),
                         "p99_latency_s": float(np.percentile(lat, 99))}
    out["served_over_plain"] = out["served"]["value"] / out["plain_loop"]["value"]
    return out


def cpu_baseline(pos, vel, col, dt, steps, threads, theta=None, multipole=0):
    from oracle import oracle
    P = oracle.params_physics(21)
    if theta is not None:
        P.theta = theta
    P.multipole = multipole
    p, v = pos, vel
    oracle.step(p, v, col, P, dt, threads=threads)  # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        p, v, _, _ = oracle.step(p, v, col, P, dt, threads=threads)
    el = time.perf_counter() - t0
    return len(pos) * steps / el, el


def accuracy(astro, ctx, pos, vel, col, P, a):
    """RMS relative force error against the FP64 direct sum (on the GPU) on a random target
    subset of the initial state, for the headline configuration (theta 0.7, monopoles) and for
    the accuracy-matched one (quadrupoles at --matched-theta), and the step rate of the
    latter."""
    n = len(pos)
    tg = np.sort(np.random.default_rng(1).choice(n, a.acc_targets, replace=False))
    ctx.upload(pos, vel, col)
    ref = ctx.direct_forces_fp64(tg)
    out = {"targets": int(a.acc_targets), "reference": "FP64 direct sum over all 500k "
           "particles (astro_direct_forces_fp64), initial state"}
    rows = {}
    for name, theta, mp, kahan in (("theta0.7_monopole", 0.7, 0, 0),
                                   ("theta0.5_monopole", 0.5, 0, 0),
                                   ("theta0.7_quadrupole", 0.7, 2, 0),
                                   ("theta0.5_quadrupole", 0.5, 2, 0),
                                   ("theta0.5_quadrupole_kahan", 0.5, 2, 1),
                                   ("theta0.55_quadrupole", 0.55, 2, 0),
                                   ("theta0.6_quadrupole", 0.6, 2, 0)):
        Q = P.copy()
        Q.theta, Q.multipole, Q.kahan = theta, mp, kahan
        ctx.set_params(Q)
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        rows[name] = rms_rel(ctx.forces()[tg], ref)
    out["rms_rel_force_error"] = rows
    # the accuracy-matched configuration, timed like the headline one
    Q = P.copy()
    Q.theta, Q.multipole = a.matched_theta, 2
    ctx.set_params(Q)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.compute_forces()
    ctx.check()
    matched_err = rms_rel(ctx.forces()[tg], ref)
    blk, chunk, _ = ctx.autotune(3)
    st = ctx.traversal_stats()
    ctx.step(a.dt, 2)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(a.dt, a.acc_steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.check()
    stage_ms = ctx.step_timed(a.dt, a.acc_steps)
    ctx.check()
    out["matched"] = {"config": "theta %g, quadrupoles, otherwise the headline workload"
                                % a.matched_theta,
                      "force_block": blk, "force_xcd_chunk": chunk,
                      "rms_rel_force_error": matched_err,
                      "value": n * a.acc_steps / wall, "unit": "particle-steps/s",
                      "ms_per_step": wall / a.acc_steps * 1e3, "steps": a.acc_steps,
                      "stage_ms": {k: round(v / a.acc_steps, 4) for k, v in stage_ms.items()},
                      "visits_per_particle": st["visits"] / n}
    ctx.set_params(P)
    return out


def bench_large(astro, a):
    """4,194,304 particles: Plummer sphere plus a 10% uniform background cube of side 20,
    Salpeter masses on [0.1, 10]; theta 0.7 and 0.5, monopoles and quadrupoles; RMS error on
    4,096 random targets against the FP64 direct sum over all particles."""
    pos, vel, col = astro.ic_plummer_background(a.large_n, 77)
    n = len(pos)
    tg = np.sort(np.random.default_rng(3).choice(n, 4096, replace=False))
    out = {"n_particles": n, "steps": a.large_steps,
           "ic": "Plummer (a = 1) + 10% uniform cube of side 20, masses power law -2.35 on "
                 "[0.1, 10], seed 77"}
    with astro.Context(n, bits_per_axis=21, emit_octree_nodes=False) as ctx:
        ctx.set_params(astro.params_physics(21))  # G and eps2 of the direct sum
        ctx.upload(pos, vel, col)
        ref = ctx.direct_forces_fp64(tg)
        for theta, mp in ((0.7, 0), (0.5, 0), (0.5, 2)):
            P = astro.params_physics(21)
            P.theta, P.multipole = theta, mp
            ctx.set_params(P)
            ctx.upload(pos, vel, col)
            ctx.build_tree()
            ctx.compute_forces()
            ctx.check()
            err = rms_rel(ctx.forces()[tg], ref)
            ctx.autotune(3)
            ctx.step(a.dt, 2)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(a.dt, a.large_steps)
            ctx.sync()
            el = time.perf_counter() - t0
            ctx.check()
            out["theta%.1f_%s" % (theta, "quadrupole" if mp else "monopole")] = {
                "value": n * a.large_steps / el, "ms_per_step": el / a.large_steps * 1e3,
                "rms_rel_force_error": err}
        out["device_bytes"] = ctx.device_bytes()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)
    import astro_amd
    pos, vel, col = astro_amd.ic_two_disks(a.n_per_disk, a.seed)
    n = len(pos)
    # the OctreeNode array is written when a caller reads it (its bytes are the same as when
    # written every step; the traversal reads the compact nodes)
    ctx = astro_amd.Context(n, bits_per_axis=21, emit_octree_nodes=astro_amd.EMIT_ON_READ)
    P = astro_amd.params_physics(21)
    if a.matched:
        P.theta, P.multipole = a.matched_theta, 2
    ctx.set_params(P)
    ctx.upload(pos, vel, col)

    ctx.build_tree()
    ctx.check()
    ctx.set_force_top_cache(a.top_cache)
    ctx.set_sort_mode(astro_amd.SORT_SEEDED if a.sort == "seeded" else astro_amd.SORT_RADIX)
    ctx.set_force_walk(1 if a.walk == "wave" else 0)
    if a.force_block and a.xcd_chunk >= 0:
        tuned_block, tuned_chunk = a.force_block, a.xcd_chunk
        ctx.set_force_block(tuned_block)
        ctx.set_force_xcd_chunk(tuned_chunk)
    else:
        tuned_block, tuned_chunk, _ = ctx.autotune(3)
    V = int(ctx.layer_counts().astype(np.int64).sum())
    st0 = ctx.traversal_stats()
    ctx.step(a.dt, max(a.warmup, 1))  # warm-up; captures the step's hipGraph
    ctx.sync()

    # timed region: K steps of astro_step (graph replays), inputs resident in HBM
    t0 = time.perf_counter()
    ctx.step(a.dt, a.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.check()
    # the next K steps as one captured graph with HIP event nodes at the stage boundaries (on
    # the library's stream): per-stage times, among them the traversal kernel's duration
    te = time.perf_counter()
    stage_ms = ctx.step_timed(a.dt, a.steps)
    ctx.sync()
    wall_events = time.perf_counter() - te
    ctx.check()
    if a.profile_only:
        if a.calibrate_copy_mb:
            calibration_copy(a.calibrate_copy_mb << 20)
        print(json.dumps({"wall_s": wall, "stage_ms": stage_ms, "force_block": tuned_block,
                          "force_xcd_chunk": tuned_chunk}))
        return
    if a.matched:
        print(json.dumps({"error": "--matched is a profiling mode; use it with --profile-only"}))
        sys.exit(2)

    st1 = ctx.traversal_stats()
    # PCIe-inclusive: host arrays in, one step, host arrays out
    p_h, v_h, c_h = ctx.download()
    ctx.sync()
    tp = time.perf_counter()
    ctx.upload(p_h, v_h, c_h)
    ctx.step(a.dt, 1)
    p_h, v_h, c_h = ctx.download()
    pcie_s = time.perf_counter() - tp
    # the same steps writing the 128-B OctreeNode array every step (beside the traversal)
    ctx.set_emit_octree_nodes(astro_amd.EMIT_EVERY_BUILD)
    ctx.step(a.dt, 2)
    ctx.sync()
    tn = time.perf_counter()
    ctx.step(a.dt, a.steps)
    ctx.sync()
    wall_every = time.perf_counter() - tn
    ctx.set_emit_octree_nodes(astro_amd.EMIT_ON_READ)
    ctx.check()

    K = a.steps
    value = n * K / wall
    force_ms = stage_ms["force"] / K
    inter = 0.5 * (st0["leaf_terms"] + st0["monopole_terms"] + st1["leaf_terms"] +
                   st1["monopole_terms"])
    achieved = FLOP_PER_INTERACTION * inter / (force_ms * 1e-3) / 1e12
    traffic = None
    valu_frac = ta_frac = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = tj.get("force_compact_bytes_per_launch")
        valu_frac = tj.get("valu_issue_frac")
        ta_frac = tj.get("ta_busy_frac")
    # the OctreeNode array is written on read only
    sb = stage_bytes(n, V, emit=False, seeded=a.sort == "seeded")
    stages = {}
    for s, ms in stage_ms.items():
        per = ms / K
        e = {"ms": round(per, 4)}
        if s in sb:
            e["gbs"] = round(sb[s] / (per * 1e-3) / 1e9, 1)
            e["hbm_frac"] = round(e["gbs"] / PEAK_HBM_GBS, 4)
        stages[s] = e

    out = {
        "metric": "particle-steps/s @500k",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-disk initial conditions)",
        "config": {"workload": "two_disks_500k", "n_particles": n, "theta": P.theta,
                   "eps2": P.eps2, "bits_per_axis": 21, "layers": 22, "dt": a.dt,
                   "fp64_top_layers": P.fp64_top_layers, "octree_nodes": "on_read",
                   "force_block": tuned_block, "force_xcd_chunk": tuned_chunk,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_force_compact", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "frac_nonpacked": achieved / PEAK_FP32_TFLOPS_NONPACKED,
                     "traffic": traffic, "kernel_ms": force_ms,
                     "valu_issue_frac": valu_frac,
                     "valu_issue_what": "SQ_INSTS_VALU x 4 cycles per wave64 instruction over "
                                        "1024 SIMDs and the kernel's busy cycles "
                                        "(GRBM_GUI_ACTIVE per XCD; rocprofv3, profiles/)",
                     "ta_busy_frac": ta_frac,
                     "limiter": "VALU issue and the texture address unit together: with "
                                "the top tree layers read from LDS a global visit costs the "
                                "CU's one address unit about as many cycles as its VALU work "
                                "costs each SIMD, and both are busy three quarters of the "
                                "kernel's cycles or more (the address unit alone bound it "
                                "while every node came from memory: busy 83%; "
                                "profiles/force_memory_pipeline_two_disks_500k.md)",
                     "interactions_per_step": inter, "flop_per_interaction": FLOP_PER_INTERACTION,
                     "visits_per_step": 0.5 * (st0["visits"] + st1["visits"])},
        "stages": stages,
        "tree_nodes": V,
        "device_bytes": ctx.device_bytes(),
        "pcie_inclusive": {"value": n / pcie_s, "ms": pcie_s * 1e3,
                           "what": "astro_upload of the host arrays + one step + astro_download"},
        "event_timed": {"ms_per_step": sum(stage_ms.values()) / K,
                        "wall_ms_per_step": wall_events / K * 1e3,
                        "what": "the K steps after the timed ones, captured as one graph with "
                                "event nodes at the stage boundaries (stages and roofline); the "
                                "wall time includes capturing and instantiating that graph"},
        "octree_nodes_every_step": {"value": n * K / wall_every,
                                    "ms_per_step": wall_every / K * 1e3,
                                    "what": "the same steps writing the 128-B OctreeNode array "
                                            "every step, beside the traversal"},
    }
    progress(f"headline: {value:.4g} particle-steps/s")
    if not a.skip_accuracy:
        progress("accuracy")
        out["accuracy"] = accuracy(astro_amd, ctx, pos, vel, col, P, a)
    ctx.close()
    if not a.skip_large:
        progress("large system")
        out["large"] = bench_large(astro_amd, a)
    if not a.skip_batched:
        progress("batched systems")
        out["batched"] = bench_batched(astro_amd, a)
    if not a.skip_serving:
        pool = serving_pool(astro_amd, a)
        out["serving"] = bench_serving(astro_amd, a, pool, a.serve_rate, a.serve_seconds, 5)
        out["serving_drain"] = bench_serving(astro_amd, a, pool, a.serve_rate, a.serve_seconds,
                                             5, drain=a.serve_drain_requests)
        out["serving_nominal"] = bench_serving(astro_amd, a, pool, a.serve_nominal_rate,
                                               a.serve_nominal_seconds, 6)
        del out["serving_nominal"]["served_over_plain"]  # both keep up: latency is the figure
    if not a.no_cpu_baseline:
        progress("CPU baseline")
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        cv, el = cpu_baseline(pos, vel, col, a.dt, a.cpu_steps, threads)
        out["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{a.cpu_steps} full oracle steps of the same 500k "
                                         f"system after one warm-up step ({el:.1f} s); sort and "
                                         f"build single-threaded as transcribed, force loop on "
                                         f"{threads} OpenMP threads"}
        out["gpu_vs_cpu"] = value / cv
        progress("CPU baseline, one thread")
        c1, el1 = cpu_baseline(pos, vel, col, a.dt, a.cpu_steps_1thread, 1)
        out["cpu_baseline"]["one_thread"] = {
            "value": c1, "cores": 1,
            "sample": f"{a.cpu_steps_1thread} full oracle steps of the same system on one "
                      f"thread after one warm-up step ({el1:.1f} s)"}
        out["gpu_vs_cpu_one_thread"] = value / c1
        if "accuracy" in out:
            cq, elq = cpu_baseline(pos, vel, col, a.dt, a.cpu_steps_matched, threads,
                                   theta=a.matched_theta, multipole=2)
            out["cpu_baseline"]["matched"] = {
                "value": cq, "sample": f"{a.cpu_steps_matched} oracle steps at theta "
                                       f"{a.matched_theta:g} "
                                       f"with quadrupoles after one warm-up step ({elq:.1f} s)"}
            out["accuracy"]["matched"]["gpu_vs_cpu"] = out["accuracy"]["matched"]["value"] / cq
    print(json.dumps(out))


if __name__ == "__main__":
    main()
