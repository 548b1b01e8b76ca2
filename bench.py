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
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (<= 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="warmup + steps only, no JSON extras (for rocprofv3 runs)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="profile mode: also copy this many MiB device-to-device (a plain copy "
                         "of known size that scales the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def stage_bytes(n, V):
    """Algorithmic HBM bytes per step of the HBM-bound stages (documented in DESIGN.md)."""
    return {
        "bbox_morton": n * (16 + 16 + 8 + 8 + 4),  # bbox read + key read/ 2 key writes + iota
        "sort": n * (24 * 8 + 8),                  # 8 digit passes of 12-B pairs + histogram
        "build": n * (4 + 16 + 8 + 4) + V * (32 + 128 + 32),
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


def cpu_baseline(pos, vel, col, dt, steps, threads):
    from oracle import oracle
    P = oracle.params_physics(21)
    p, v = pos, vel
    oracle.step(p, v, col, P, dt, threads=threads)  # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        p, v, _, _ = oracle.step(p, v, col, P, dt, threads=threads)
    el = time.perf_counter() - t0
    return len(pos) * steps / el, el


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)
    import astro_amd
    pos, vel, col = astro_amd.ic_two_disks(a.n_per_disk, a.seed)
    n = len(pos)
    ctx = astro_amd.Context(n, bits_per_axis=21, emit_octree_nodes=True)
    P = astro_amd.params_physics(21)
    ctx.set_params(P)
    ctx.upload(pos, vel, col)

    ctx.build_tree()
    ctx.check()
    V = int(ctx.layer_counts().astype(np.int64).sum())
    st0 = ctx.traversal_stats()
    ctx.step_timed(a.dt, max(a.warmup, 1))
    ctx.sync()

    t0 = time.perf_counter()
    stage_ms = ctx.step_timed(a.dt, a.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.check()
    if a.profile_only:
        if a.calibrate_copy_mb:
            calibration_copy(a.calibrate_copy_mb << 20)
        print(json.dumps({"wall_s": wall, "stage_ms": stage_ms}))
        return

    st1 = ctx.traversal_stats()
    # graph-replayed steps of the same state, for reference
    ctx.set_use_graph(True)
    ctx.step(a.dt, 2)
    ctx.sync()
    tg = time.perf_counter()
    ctx.step(a.dt, a.steps)
    ctx.sync()
    wall_graph = time.perf_counter() - tg
    ctx.check()

    # the same steps without writing the 128-B OctreeNode array (compact nodes only)
    ctx.set_emit_octree_nodes(False)
    ctx.step(a.dt, 2)
    ctx.sync()
    tn = time.perf_counter()
    ctx.step(a.dt, a.steps)
    ctx.sync()
    wall_noemit = time.perf_counter() - tn
    ctx.set_emit_octree_nodes(True)
    ctx.check()

    K = a.steps
    value = n * K / wall
    force_ms = stage_ms["force"] / K
    inter = 0.5 * (st0["leaf_terms"] + st0["monopole_terms"] + st1["leaf_terms"] +
                   st1["monopole_terms"])
    achieved = FLOP_PER_INTERACTION * inter / (force_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("force_compact_bytes_per_launch")
    sb = stage_bytes(n, V)
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
                   "fp64_top_layers": P.fp64_top_layers, "octree_nodes_emitted": True,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_force_compact", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "frac_nonpacked": achieved / PEAK_FP32_TFLOPS_NONPACKED,
                     "traffic": traffic, "kernel_ms": force_ms,
                     "interactions_per_step": inter, "flop_per_interaction": FLOP_PER_INTERACTION,
                     "visits_per_step": 0.5 * (st0["visits"] + st1["visits"])},
        "stages": stages,
        "tree_nodes": V,
        "graph_replay": {"value": n * K / wall_graph, "ms_per_step": wall_graph / K * 1e3},
        "without_octree_nodes": {"value": n * K / wall_noemit,
                                 "ms_per_step": wall_noemit / K * 1e3},
    }
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        cv, el = cpu_baseline(pos, vel, col, a.dt, a.cpu_steps, threads)
        out["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{a.cpu_steps} full oracle steps of the same 500k "
                                         f"system after one warm-up step ({el:.1f} s); sort and "
                                         f"build single-threaded as transcribed, force loop on "
                                         f"{threads} OpenMP threads"}
        out["gpu_vs_cpu"] = value / cv
    print(json.dumps(out))


if __name__ == "__main__":
    main()
