"""GPU parity at the benchmark configurations' own sizes.

* the large single system: 4,194,304 particles (Plummer sphere + 10% uniform cube of side 20,
  power-law masses), theta 0.5 and 0.7;
* the batched form: 256 Plummer systems of 1,024 .. 65,536 particles;
* the serving form: requests drawn from the serving stream's own distributions (Plummer or
  two-disk systems, N log-uniform in [2,048, 2^20], 10 .. 1,000 steps);
* the 500k two-disk system over 20 steps.

At full size the tests check what does not depend on the oracle finishing (sortedness, the
permutation against numpy's stable argsort, layer counts against the distinct key prefixes),
then compare the whole tree and a seeded target subset with the oracle bit for bit.  Each
test prints its case before touching the GPU."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

from test_gpu_batched import make_batch, run_server

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def layer_counts_from_keys(keys, layers):
    """Valid nodes per layer: the distinct key prefixes (keys >> 3l) of the sorted keys."""
    out = []
    for l in range(layers):
        k = keys >> np.uint64(3 * l)
        out.append(int(1 + np.count_nonzero(k[1:] != k[:-1])) if len(k) else 0)
    return out


def oracle_forces_at(orc, tree, pos, col, P, targets):
    f = np.zeros((len(targets), 3), np.float32)
    for a, i in enumerate(targets):
        w, _ = orc.barnes_hut(tree, pos, col, P, targets=(int(i), int(i) + 1))
        f[a] = w[i, :3]
    return f


@pytest.fixture(scope="module")
def large_system(astro):
    return astro.ic_plummer_background(4_194_304, 77)


def test_large_plummer_background_tree_and_forces(astro, orc, large_system):
    # 4,194,304 particles: keys, permutation, layer counts and the whole OctreeNode array
    # (7.8 GB) at full size; forces of 4,096 seeded targets bit-exact against the oracle at
    # theta 0.7 and 0.5, with their RMS error against the FP64 direct sum; the integrator's
    # step of those targets
    pos, vel, col = large_system
    n = len(pos)
    Po = orc.params_physics(21)
    t = orc.build_tree(pos, Po)
    tg = np.sort(np.random.default_rng(1).choice(n, 4096, replace=False))
    with astro.Context(n, bits_per_axis=21) as ctx:
        for theta, rms_bound in ((0.7, 1e-2), (0.5, 5e-3)):
            say("large 4,194,304 Plummer + cube, theta", theta)
            P = astro.params_physics(21)
            P.theta = Po.theta = theta
            ctx.set_params(P)
            ctx.upload(pos, vel, col)
            ctx.build_tree()
            ctx.compute_forces()
            ctx.check()
            if theta == 0.7:  # the tree does not depend on theta
                keys = ctx.read("SORTED_CODES", np.uint64, n)
                perm = ctx.read("SORTED_INDICES", np.uint32, n)
                assert np.array_equal(ctx.read("MORTON_CODES", np.uint64, n), t["codes"])
                assert np.all(keys[1:] >= keys[:-1])
                o = np.argsort(t["codes"], kind="stable").astype(np.uint32)
                assert np.array_equal(perm, o)
                counts = list(ctx.layer_counts())
                assert counts == layer_counts_from_keys(keys, 22) == t["count"]
                assert same_bytes(ctx.octree_nodes(), t["nodes"])
            f = ctx.forces()
            want = oracle_forces_at(orc, t, pos, col, Po, tg)
            assert np.array_equal(f[tg, :3], want), f"theta {theta}: forces differ"
            ref = ctx.direct_forces_fp64(tg)
            d = f[tg, :3].astype(np.float64) - ref
            rms = float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))
            assert rms < rms_bound, (theta, rms)
            # one integrator step of the targets (BarnesHut.metal:169-196)
            ctx.update_spheres(np.float32(1e-3))
            p2, v2, _ = ctx.download()
            f4 = np.zeros((n, 4), np.float32)
            f4[tg, :3] = want
            wp, wv = orc.update_spheres(pos[tg], vel[tg], f4[tg], col[tg], np.float32(1e-3), Po)
            assert np.array_equal(p2["position"][tg], wp["position"])
            assert np.array_equal(v2["velocity"][tg], wv["velocity"])


@pytest.mark.parametrize("kahan", [0, 1])
def test_batched_256_systems_full_sizes(astro, orc, kahan):
    # 256 Plummer systems, N log-uniform in [1,024, 65,536] (seed 7, 4.0 M particles), three
    # batched graph steps: 8 systems spread over the size range bit-exact against the
    # per-system oracle; every system's segment sorted in place, one root per system, every
    # position finite
    say("batched 256 systems [1024, 65536], kahan", kahan)
    sizes = astro.ic_batch_sizes(256, 1024, 65536, 7)
    parts, pos, vel, col, off = make_batch(astro, sizes, seed0=2000)
    P, Po = astro.params_physics(21), orc.params_physics(21)
    P.kahan = Po.kahan = kahan
    steps = 3
    with astro.Context(len(pos), bits_per_axis=21, max_systems=256) as ctx:
        ctx.set_params(P)
        ctx.upload_systems(pos, vel, col, off)
        ctx.build_tree()
        ctx.check()
        assert ctx.layer_counts()[-1] == 256
        keys = ctx.read("SORTED_CODES", np.uint64, len(pos))
        perm = ctx.read("SORTED_INDICES", np.uint32, len(pos))
        for k in range(256):
            a, b = off[k], off[k + 1]
            assert np.all((perm[a:b] >= a) & (perm[a:b] < b)), k
            assert np.all(keys[a + 1:b] >= keys[a:b - 1]), k
        ctx.step(1e-3, steps)
        ctx.check()
        pb, vb, _ = ctx.download()
    assert np.isfinite(pb["position"]).all() and np.isfinite(vb["velocity"]).all()
    order = np.argsort(sizes, kind="stable")
    pick = order[np.linspace(0, 255, 8).astype(int)]
    for k in pick:
        p, v, c = parts[k]
        for _ in range(steps):
            p, v, _, _ = orc.step(p, v, c, Po, 1e-3, threads=16)
        sl = slice(off[k], off[k + 1])
        assert np.array_equal(pb[sl]["position"], p["position"]), f"system {k} ({sizes[k]})"
        assert np.array_equal(vb[sl]["velocity"], v["velocity"]), f"system {k} ({sizes[k]})"


def serving_requests(astro, count, seed):
    """Requests from the serving stream's distributions: Plummer or two-disk (alternating),
    N log-uniform in [2,048, 2^20], steps uniform in [10, 1,000], arrival ticks spread."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(2048), np.log(1 << 20))))
        steps = int(rng.integers(10, 1001))
        if k % 2:
            p, v, c = astro.ic_two_disks(max(1, n // 2), 900 + k)
        else:
            p, v, c = astro.ic_plummer(n, 900 + k)
        reqs.append((100 + k, int(rng.integers(0, 40)), p, v, c, steps))
    return reqs


def test_serving_stream_requests_alone_vs_mix(astro):
    # 12 requests of the serving stream's own distributions through the continuous-batching
    # server: every result equals the same request run alone (plain single-system steps)
    say("serving stream: 12 requests alone vs mix")
    P = astro.params_physics(21)
    reqs = serving_requests(astro, 12, 4242)
    total = sum(len(r[2]) for r in reqs)
    mix, recs = run_server(astro, P, reqs, total, 16)
    assert len(recs) == 12
    for rid, _, p, v, c, steps in reqs:
        with astro.Context(len(p), bits_per_axis=21, emit_octree_nodes=False) as ctx:
            ctx.set_params(P)
            ctx.upload(p, v, c)
            ctx.step(1e-3, steps)
            ctx.check()
            pa, va, _ = ctx.download()
        assert same_bytes(mix[rid][0], pa), f"request {rid} (N {len(p)}, {steps} steps)"
        assert same_bytes(mix[rid][1], va)


def test_two_disks_500k_twenty_steps(astro, orc):
    # the 500,000-particle two-disk system, 20 graph steps bit-exact against 20 oracle steps
    say("two disks 500k, 20 steps")
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    with astro.Context(len(pos), bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        ctx.step(1e-3, 20)
        ctx.check()
        pg, vg, _ = ctx.download()
    Po = orc.params_physics(21)
    p, v = pos, vel
    for _ in range(20):
        p, v, _, _ = orc.step(p, v, col, Po, 1e-3, threads=16)
    assert np.array_equal(pg["position"], p["position"])
    assert np.array_equal(vg["velocity"], v["velocity"])
