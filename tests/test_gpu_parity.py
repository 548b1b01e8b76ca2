"""GPU parity: libastro's HIP path against the CPU oracle on the same seeded inputs.

Integer work (keys, permutation, topology) must be bit-exact; forces and positions must be
within 1e-5 relative (norm-wise, per particle) and, because kernel and oracle perform the same
IEEE operations in the same order, are also checked for bit equality.  Each test prints its
case before touching the GPU."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

from astro_amd.structs import COLOR_TYPE, INVALID_MORTON_CODE, OCTREE_NODE, POSITION_MASS

pytestmark = pytest.mark.gpu
FORCE_RTOL = 1e-5


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    den = np.maximum(np.linalg.norm(b, axis=1), 1e-30)
    return np.linalg.norm(a - b, axis=1) / den


@pytest.fixture(scope="module")
def small_ctx(astro):
    c = astro.Context(70000, bits_per_axis=21, max_systems=1, emit_octree_nodes=True)
    yield c
    c.close()


def plummer(astro, n, seed):
    return astro.ic_plummer(n, seed)


CASES = [
    ("compat_205_of_2000", "compat", 2000, 205),
    ("physics_B7_16384", 7, 16384, None),
    ("physics_B21_16384", 21, 16384, None),
    ("physics_B21_3001", 21, 3001, None),
    ("physics_B6_4000", 6, 4000, None),
]


def setup_case(astro, orc, ctx, case):
    name, prof, n, nt = case
    if prof == "compat":
        rng = np.random.default_rng(5)
        pos = np.zeros(n, POSITION_MASS)
        pos["position"] = rng.uniform(-2e6, 2e6, (n, 3)).astype(np.float32)
        pos["mass"] = np.where(np.arange(n) < 5, 1e4, np.where(np.arange(n) < nt, 1e3, 1.0))
        col = np.zeros(n, COLOR_TYPE)
        col["type"] = np.where(np.arange(n) < 5, 0, np.where(np.arange(n) < nt, 1, 2))
        vel = np.zeros(n, astro.VELOCITY_RADIUS)
        vel["velocity"] = rng.uniform(-2e4, 2e4, (n, 3)).astype(np.float32)
        P, Po = astro.params_compat(), orc.params_compat()
    else:
        pos, vel, col = plummer(astro, n, 42)
        P, Po = astro.params_physics(prof), orc.params_physics(prof)
    ctx.set_params(P)
    ctx.upload(pos, vel, col, n_tree=nt or 0)
    return pos, vel, col, Po, (nt or n)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fused_build_bitexact(astro, orc, small_ctx, case):
    say("fused build", case[0])
    ctx = small_ctx
    pos, vel, col, Po, nt = setup_case(astro, orc, ctx, case)
    ctx.set_emit_octree_nodes(True)
    ctx.build_tree()
    ctx.check()
    t = orc.build_tree(pos, Po, n_tree=nt)
    assert np.array_equal(ctx.read("MORTON_CODES", np.uint64, nt), t["codes"])
    assert np.array_equal(ctx.read("SORTED_CODES", np.uint64, nt), t["sorted_codes"])
    assert np.array_equal(ctx.read("SORTED_INDICES", np.uint32, nt), t["sorted_idx"])
    if Po.quantizer == 1:
        assert np.array_equal(ctx.read("CUBES", np.float32, 4), t["cube"])
    assert list(ctx.layer_counts()) == t["count"]
    got = ctx.octree_nodes()
    assert len(got) == len(t["nodes"])
    bad = np.nonzero(got.view(np.uint8).reshape(-1, 128) != t["nodes"].view(np.uint8).reshape(-1, 128))[0]
    assert len(bad) == 0, f"{len(np.unique(bad))} nodes differ, first {np.unique(bad)[:5]}"


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_forces_bitexact(astro, orc, small_ctx, case):
    say("forces", case[0])
    ctx = small_ctx
    pos, vel, col, Po, nt = setup_case(astro, orc, ctx, case)
    ctx.build_tree()
    ctx.compute_forces()
    ctx.check()
    t = orc.build_tree(pos, Po, n_tree=nt)
    want, _ = orc.barnes_hut(t, pos, col, Po, n_tree=nt, threads=8)
    got = ctx.forces()
    e = rel_err(got, want)
    assert e.max() <= FORCE_RTOL, f"max rel err {e.max():.3e}"
    assert np.array_equal(got[:, :3], want[:, :3]), f"not bit-exact, max rel {e.max():.3e}"


@pytest.mark.parametrize("kahan", [0, 1])
def test_forces_kahan(astro, orc, small_ctx, kahan):
    say("forces kahan", kahan)
    ctx = small_ctx
    pos, vel, col = plummer(astro, 16384, 7)
    P, Po = astro.params_physics(21), orc.params_physics(21)
    P.kahan = Po.kahan = kahan
    ctx.set_params(P)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.compute_forces()
    t = orc.build_tree(pos, Po)
    want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
    assert np.array_equal(ctx.forces()[:, :3], want[:, :3])


def run_stage_sequence(ctx, nt, layers_cap_off):
    """Renderer.swift:511-534 through the stage calls."""
    cap, off = layers_cap_off
    total = int(off[-1] + cap[-1])
    ctx.clear_buffer("SORTED_CODES", nt, 8)
    ctx.clear_buffer("SORTED_INDICES", nt, 4)
    ctx.clear_octree_node_buffer(total)
    ctx.generate_morton_codes(nt)
    ctx.fill_indices("UNSORTED_INDICES", 0, nt)
    ctx.write("PARENT_COUNT", np.array([nt], np.uint32))
    ctx.radix_sort("UNSORTED_CODES", "UNSORTED_INDICES", "SORTED_CODES", "SORTED_INDICES", nt,
                   "PARENT_COUNT")
    ctx.count_unique_morton_codes(False, nt)
    ctx.aggregate_leaf_nodes(nt)
    ctx.build_internal_nodes(nt)


@pytest.mark.parametrize("mode", ["every_build", "on_read"])
@pytest.mark.parametrize("graph", [True, False])
def test_step_writes_octree_nodes(astro, orc, graph, mode):
    # every build: a step writes the OctreeNode array on a second stream beside the traversal
    # and joins it; on read: the array of the last built tree is written when it is read.
    # Either way, after two steps the array must equal the oracle's tree of the positions the
    # second step started from, stale slots of the first step's tree emptied
    say("octree nodes written by steps, graph", graph, "mode", mode)
    pos, vel, col = astro.ic_plummer(20000, 31)
    em = astro.EMIT_EVERY_BUILD if mode == "every_build" else astro.EMIT_ON_READ
    with astro.Context(20000, bits_per_axis=21, emit_octree_nodes=em) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.set_use_graph(graph)
        ctx.upload(pos, vel, col)
        ctx.step(1e-3, 1)
        ctx.check()
        p1, _, _ = ctx.download()
        if mode == "on_read":  # read between the steps too: the tree the first step built
            t0 = orc.build_tree(pos, orc.params_physics(21))
            assert same_bytes(ctx.octree_nodes(), t0["nodes"])
        ctx.step(1e-3, 3)  # two more builds before the next read (on read: nothing written)
        ctx.check()
        p3, _, _ = ctx.download()
        ctx.step(1e-3, 1)
        ctx.check()
        t = orc.build_tree(p3, orc.params_physics(21))
        assert same_bytes(ctx.octree_nodes(), t["nodes"])
        assert p1.tobytes() != p3.tobytes()


@pytest.mark.parametrize("case", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_stage_path_matches_oracle(astro, orc, small_ctx, case):
    say("stage path", case[0])
    ctx = small_ctx
    pos, vel, col, Po, nt = setup_case(astro, orc, ctx, case)
    run_stage_sequence(ctx, nt, ctx.layer_table())
    ctx.barnes_hut()
    ctx.check()
    t = orc.build_tree(pos, Po, n_tree=nt)
    got = ctx.octree_nodes()
    assert same_bytes(got, t["nodes"])
    want, _ = orc.barnes_hut(t, pos, col, Po, n_tree=nt, threads=8)
    assert np.array_equal(ctx.forces()[:, :3], want[:, :3])
    # and the integrator stage (BarnesHut.metal:169-196)
    ctx.update_spheres(np.float32(1 / 30))
    p2, v2, _ = ctx.download()
    wp, wv = orc.update_spheres(pos, vel, want, col, np.float32(1 / 30), Po)
    assert np.array_equal(p2["position"], wp["position"])
    assert np.array_equal(v2["velocity"], wv["velocity"])


@pytest.mark.parametrize("graph", [True, False])
def test_ten_steps_bitexact(astro, orc, small_ctx, graph):
    # 16,384 Plummer, B=7 (21-bit keys), theta 0.7, eps^2 0.01, dt 1e-3, 10 steps
    say("10 steps, graph", graph)
    ctx = small_ctx
    pos, vel, col = plummer(astro, 16384, 42)
    P, Po = astro.params_physics(7), orc.params_physics(7)
    ctx.set_params(P)
    ctx.set_use_graph(graph)
    ctx.upload(pos, vel, col)
    ctx.step(1e-3, 10)
    ctx.check()
    p2, v2, _ = ctx.download()
    p, v = pos, vel
    for _ in range(10):
        p, v, f, _ = orc.step(p, v, col, Po, 1e-3, threads=8)
    ctx.set_use_graph(True)
    assert np.array_equal(p2["position"], p["position"])
    assert np.array_equal(v2["velocity"], v["velocity"])


def test_graph_replays_match_eager(astro, small_ctx):
    say("graph replays vs eager")
    ctx = small_ctx
    pos, vel, col = plummer(astro, 20000, 3)
    ctx.set_params(astro.params_physics(21))
    outs = []
    for graph in (False, True):
        ctx.set_use_graph(graph)
        ctx.upload(pos, vel, col)
        for _ in range(4):  # replays with host synchronisation between them
            ctx.step(1e-3, 1)
            ctx.sync()
        outs.append(ctx.download()[0].tobytes())
    ctx.set_use_graph(True)
    assert outs[0] == outs[1]


@pytest.mark.parametrize("n", [1, 5, 2047, 2048, 2049, 300000])
def test_scan(astro, small_ctx, n):
    say("scan", n)
    rng = np.random.default_rng(n)
    ctx = astro.Context(max(n, 16), bits_per_axis=7) if n > 70000 else small_ctx
    a = rng.integers(0, 5, n).astype(np.uint32)
    ctx.write("FLAGS", a)
    ctx.scan("FLAGS", "SCAN", n)
    ctx.check()
    got = ctx.read("SCAN", np.uint32, n)
    assert np.array_equal(got, np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.uint32))


@pytest.mark.parametrize("bit", [0, 5, 63])
def test_split_pass(astro, orc, small_ctx, bit):
    say("split pass", bit)
    ctx = small_ctx
    rng = np.random.default_rng(bit)
    n, count = 5000, 4321
    k = rng.integers(0, 1 << 63, n, dtype=np.uint64) | (rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63))
    idx = np.arange(n, dtype=np.uint32)
    ctx.write("UNSORTED_CODES", k)
    ctx.write("UNSORTED_INDICES", idx)
    ctx.write("PARENT_COUNT", np.array([count], np.uint32))
    ctx.split_pass("UNSORTED_CODES", "UNSORTED_INDICES", "SORTED_CODES", "SORTED_INDICES", bit, n,
                   "PARENT_COUNT")
    ctx.check()
    ko, io = orc.split_pass(k[:count], idx[:count], bit)
    assert np.array_equal(ctx.read("SORTED_CODES", np.uint64, count), ko)
    assert np.array_equal(ctx.read("SORTED_INDICES", np.uint32, count), io)


@pytest.mark.parametrize("kind", ["random64", "invalid_tail", "equal", "few", "small"])
def test_radix_sort_generic(astro, small_ctx, kind):
    say("radix sort", kind)
    ctx = small_ctx
    rng = np.random.default_rng(11)
    n = 60000 if kind != "small" else 37
    count = n - 123 if kind != "small" else n
    if kind == "random64":
        k = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n).astype(np.uint64)
    elif kind == "invalid_tail":
        k = rng.integers(0, 1 << 40, n, dtype=np.uint64)
        k[rng.integers(0, n, 3000)] = INVALID_MORTON_CODE
    elif kind == "equal":
        k = np.full(n, 77, np.uint64)
    elif kind == "few":
        k = rng.integers(0, 3, n).astype(np.uint64) << np.uint64(61)
    else:
        k = rng.integers(0, 10, n).astype(np.uint64)
    idx = rng.permutation(n).astype(np.uint32)
    ctx.write("UNSORTED_CODES", k)
    ctx.write("UNSORTED_INDICES", idx)
    ctx.write("PARENT_COUNT", np.array([count], np.uint32))
    ctx.radix_sort("UNSORTED_CODES", "UNSORTED_INDICES", "SORTED_CODES", "SORTED_INDICES", n,
                   "PARENT_COUNT")
    ctx.check()
    o = np.argsort(k[:count], kind="stable")
    assert np.array_equal(ctx.read("SORTED_CODES", np.uint64, count), k[:count][o])
    assert np.array_equal(ctx.read("SORTED_INDICES", np.uint32, count), idx[:count][o])


def test_radix_sort_many_tiles(astro):
    # 3 M keys (2,930 tiles of 1,024) with many repeats (ties keep their input order), twice
    # in a row with different live counts, against numpy's stable argsort
    say("radix sort, many tiles")
    n = 3_000_000
    rng = np.random.default_rng(21)
    with astro.Context(n, bits_per_axis=21) as ctx:
        for rep in range(2):
            k = (rng.integers(0, 1 << 20, n).astype(np.uint64) << np.uint64(40)) | \
                rng.integers(0, 4, n).astype(np.uint64)
            count = n - 1000 * rep
            idx = rng.permutation(n).astype(np.uint32)
            ctx.write("UNSORTED_CODES", k)
            ctx.write("UNSORTED_INDICES", idx)
            ctx.write("PARENT_COUNT", np.array([count], np.uint32))
            ctx.radix_sort("UNSORTED_CODES", "UNSORTED_INDICES", "SORTED_CODES", "SORTED_INDICES",
                           n, "PARENT_COUNT")
            ctx.check()
            o = np.argsort(k[:count], kind="stable")
            assert np.array_equal(ctx.read("SORTED_CODES", np.uint64, count), k[:count][o])
            assert np.array_equal(ctx.read("SORTED_INDICES", np.uint32, count), idx[:count][o])


@pytest.mark.parametrize("multipole", [0, 2])
def test_tree_past_2_28_nodes(astro, multipole):
    # 21 M particles at B = 21: the layer table holds more than 2^28 nodes, so the byte
    # offsets of the 16-B node halves pass 4 GiB and the traversal takes its 64-bit
    # addressing (with quadrupoles: links from the link array, none in the moment rows); the
    # forces of 512 random targets must match the FP64 direct sum to the tree's usual accuracy
    say("tree of more than 2^28 nodes, multipole", multipole)
    n = 21_000_000
    pos, vel, col = astro.ic_plummer(n, 9)
    with astro.Context(n, bits_per_axis=21, emit_octree_nodes=False) as ctx:
        P = astro.params_physics(21)
        P.multipole = multipole
        ctx.set_params(P)
        ctx.upload(pos, vel, col)
        cap, off = ctx.layer_table()
        assert int(off[-1] + cap[-1]) >= 1 << 28
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        tg = np.sort(np.random.default_rng(3).choice(n, 512, replace=False))
        ref = ctx.direct_forces_fp64(tg)
        d = ctx.forces()[tg, :3].astype(np.float64) - ref
        rms = float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))
        assert rms < (2e-2 if multipole == 0 else 5e-3), rms
        if multipole == 2:  # no links in the moment rows of such a tree
            cap, off = ctx.layer_table()
            cnt = ctx.layer_counts()
            for l in (0, 11, len(cap) - 1):
                m = min(int(cnt[l]), 4096)
                q = ctx.read("TREE_MOMENTS", np.uint32, 8 * m, offset=8 * int(off[l]))
                assert not q.reshape(m, 8)[:, 6:].any(), l


@pytest.mark.parametrize("multipole,theta", [(0, 0.7), (2, 0.5), (2, 0.55)])
def test_two_disks_500k(astro, orc, multipole, theta):
    # the 500,000-particle two-disk system at 63-bit keys: whole tree bit-exact, forces on a
    # 4,096-target subset bit-exact, sortedness and permutation at full size; theta 0.55 with
    # quadrupoles is the accuracy-matched setting (RMS force error <= 1e-3)
    say("two disks 500k, multipole", multipole, "theta", theta)
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    n = len(pos)
    with astro.Context(n, bits_per_axis=21, emit_octree_nodes=True) as ctx:
        P, Po = astro.params_physics(21), orc.params_physics(21)
        P.multipole = Po.multipole = multipole
        P.theta = Po.theta = theta
        ctx.set_params(P)
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        keys = ctx.read("SORTED_CODES", np.uint64, n)
        perm = ctx.read("SORTED_INDICES", np.uint32, n)
        assert np.all(keys[1:] >= keys[:-1])
        assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
        t = orc.build_tree(pos, Po)
        assert np.array_equal(perm, t["sorted_idx"])
        assert list(ctx.layer_counts()) == t["count"]
        assert same_bytes(ctx.octree_nodes(), t["nodes"])
        f = ctx.forces()
        tg = np.random.default_rng(1).choice(n, 4096, replace=False)
        tg.sort()
        ok = 0
        for a in range(0, 4096, 512):  # oracle on contiguous runs of targets
            for i in tg[a:a + 512]:
                w, _ = orc.barnes_hut(t, pos, col, Po, targets=(int(i), int(i) + 1))
                ok += np.array_equal(f[i, :3], w[i, :3])
        assert ok == 4096


def test_forces_need_a_built_tree(astro):
    # the traversal trusts the tree's child indices, so the library refuses to traverse before
    # a tree is built for the uploaded particles (and after a change of the layer table)
    say("traversal refused without a tree")
    pos, vel, col = astro.ic_plummer(1000, 3)
    with astro.Context(1000, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        with pytest.raises(astro.AstroError):
            ctx.compute_forces()
        ctx.build_tree()
        ctx.compute_forces()
        ctx.upload(pos[:500], vel[:500], col[:500])
        with pytest.raises(astro.AstroError):
            ctx.compute_forces()
        ctx.check()


def test_astro_check_reports_clean(astro, small_ctx):
    say("error word clean")
    small_ctx.check()


# ------------------------------------------------- quadrupole option (multipole = 2)
QUAD_CASES = [
    # name, bits, n, fp64_top_layers, kahan
    ("B21_16384", 21, 16384, 4, 0),
    ("B21_16384_fp32", 21, 16384, 0, 0),
    ("B7_16384_fp64", 7, 16384, 99, 0),
    ("B21_16384_kahan", 21, 16384, 4, 1),
    ("B5_3001", 5, 3001, 4, 0),
]


def quad_params(astro, orc, bits, fp64, kahan, theta=0.5):
    P, Po = astro.params_physics(bits), orc.params_physics(bits)
    for q in (P, Po):
        q.multipole, q.fp64_top_layers, q.kahan, q.theta = 2, fp64, kahan, theta
    return P, Po


@pytest.mark.parametrize("case", QUAD_CASES, ids=[c[0] for c in QUAD_CASES])
def test_quadrupole_bitexact(astro, orc, small_ctx, case):
    # second moments (in the emitted nodes' emittedColor/emittedColorCenter slots) and the
    # quadrupole forces equal the oracle's bit for bit
    name, bits, n, fp64, kahan = case
    say("quadrupole", name)
    ctx = small_ctx
    pos, vel, col = plummer(astro, n, 9)
    P, Po = quad_params(astro, orc, bits, fp64, kahan)
    ctx.set_params(P)
    ctx.set_emit_octree_nodes(True)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.compute_forces()
    ctx.check()
    t = orc.build_tree(pos, Po)
    got = ctx.octree_nodes()
    assert same_bytes(got, t["nodes"])
    assert got["emittedColor"].any()
    want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
    f = ctx.forces()
    e = rel_err(f, want)
    assert e.max() <= FORCE_RTOL, f"max rel err {e.max():.3e}"
    assert np.array_equal(f[:, :3], want[:, :3])


def test_quadrupole_steps_and_back_to_monopole(astro, orc, small_ctx):
    say("quadrupole 5 graph steps, then monopole")
    ctx = small_ctx
    pos, vel, col = plummer(astro, 20000, 13)
    P, Po = quad_params(astro, orc, 21, 4, 0)
    ctx.set_params(P)
    ctx.set_use_graph(True)
    ctx.upload(pos, vel, col)
    ctx.step(1e-3, 5)
    ctx.check()
    p2, v2, _ = ctx.download()
    p, v = pos, vel
    for _ in range(5):
        p, v, _, _ = orc.step(p, v, col, Po, 1e-3, threads=8)
    assert np.array_equal(p2["position"], p["position"])
    assert np.array_equal(v2["velocity"], v["velocity"])
    # the stage traversal reads no second moments and refuses the option
    with pytest.raises(astro.AstroError):
        ctx.barnes_hut()
    # switching back restores the reference's monopole tree (zero emitted slots)
    P.multipole = Po.multipole = 0
    ctx.set_params(P)
    ctx.set_emit_octree_nodes(True)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.compute_forces()
    t = orc.build_tree(pos, Po)
    assert same_bytes(ctx.octree_nodes(), t["nodes"])
    want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
    assert np.array_equal(ctx.forces()[:, :3], want[:, :3])


def test_fast_division_matches_ieee(astro, small_ctx):
    # the traversal's division without operand scaling, on 2^26 pairs of its domain
    say("fast division self-check")
    assert small_ctx.check_division(1 << 26, seed=12345) == 0


def test_square_root_exhaustive(astro, small_ctx):
    # the fast paths' square roots (hardware estimate + one-ulp correction; reciprocal square
    # root + one Goldschmidt step) on every float in [2^-96, FLT_MAX] against the double square
    # root rounded to float
    say("exhaustive square-root self-check")
    n, bad = small_ctx.check_sqrt()
    assert n == (0x7F800000 - ((127 - 96) << 23)) and bad == 0


def test_reciprocal_exhaustive(astro, small_ctx):
    # the fast division's premise (Markstein): the refined hardware reciprocal is correctly
    # rounded on every float in [2^-45, 2^46)
    say("exhaustive reciprocal self-check")
    n, bad = small_ctx.check_reciprocal()
    assert n == ((127 + 46) << 23) - ((127 - 45) << 23) and bad == 0


def test_squared_opening_test_matches_sqrt(astro, small_ctx):
    # with quadrupoles and the fast division the traversal opens a node unless r2 >= c0 4^l,
    # in place of RN(sqrt(r2)) >= rmin0 2^l; 2^24 cases within 8 ulps of the threshold or near
    say("squared opening test self-check")
    assert small_ctx.check_open_test(1 << 24, seed=777) == 0


@pytest.mark.parametrize("scale", [1.0, 3.0e4])
def test_forces_bitexact_both_division_paths(astro, orc, small_ctx, scale):
    # scale 1: cubes inside +-8192, masses bounded: the fast division; scale 3e4: cubes
    # beyond it, the full IEEE division -- both equal the oracle bit for bit
    say("division path, scale", scale)
    ctx = small_ctx
    pos, vel, col = plummer(astro, 16384, 21)
    pos = pos.copy()
    pos["position"] *= np.float32(scale)
    P, Po = astro.params_physics(21), orc.params_physics(21)
    ctx.set_params(P)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.compute_forces()
    ctx.check()
    t = orc.build_tree(pos, Po)
    want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
    assert np.array_equal(ctx.forces()[:, :3], want[:, :3])


def test_lighting_stage_and_step(astro, orc, small_ctx):
    # BarnesHut.metal:103-167 through the stage call and inside the compat step (5 stars)
    say("lighting pass, compat 2000/205")
    ctx = small_ctx
    pos, vel, col, Po, nt = setup_case(astro, orc, ctx, CASES[0])
    col = col.copy()
    col["color"] = np.random.default_rng(8).uniform(0, 1, (len(col), 4)).astype(np.float32)
    ctx.upload(pos, vel, col, n_tree=nt)
    ctx.lighting_pass(5)
    _, _, got = ctx.download()
    want = orc.lighting_pass(pos, col, 5)
    assert np.array_equal(got["color"], want["color"]) and np.array_equal(got["type"], want["type"])
    ctx.set_use_graph(True)
    ctx.upload(pos, vel, col, n_tree=nt)
    ctx.step(np.float32(1 / 60), 3)
    ctx.check()
    p2, v2, c2 = ctx.download()
    p, v, c = pos, vel, col
    for _ in range(3):
        p, v, c, _, _ = orc.step(p, v, c, Po, np.float32(1 / 60), n_tree=nt, threads=8,
                                 return_col=True)
    assert np.array_equal(c2["color"], c["color"])
    assert np.array_equal(p2["position"], p["position"])


EDGE_CASES = ["one", "two", "all_equal", "pairs_equal", "massless_half", "b1_tree", "line",
              "equal_run_3000", "dense_cluster", "equal_run_3000_bg", "dense_cluster_bg",
              "dense_cluster_bg_fp64"]


def longest_run_without_head(sorted_keys, layer):
    """Positions of the largest node of `layer` (the longest run of equal keys >> 3 layer)."""
    k = sorted_keys >> np.uint64(3 * layer)
    heads = np.concatenate([[0], np.nonzero(k[1:] != k[:-1])[0] + 1, [len(k)]])
    return int(np.diff(heads).max())


@pytest.mark.parametrize("multipole", [0, 2])
@pytest.mark.parametrize("kind", EDGE_CASES)
def test_edge_cases_bitexact(astro, orc, small_ctx, kind, multipole):
    # degenerate inputs through the fused build, traversal and 3 steps, against the oracle:
    # one and two particles, every particle on one point (one leaf), coincident pairs
    # (key collisions), massless particles, the smallest tree (B = 1), a 1-D line; a leaf of
    # 3,000 equal keys and a cluster of 5,000 particles inside one node of layer 11, so that
    # nodes of the lower layers run past a 1,024-position tile and its halo; with 40,000
    # background particles (more than 32,768, so that the fused lower-layer build runs and
    # its leftover pass builds those nodes), once with every layer summed in FP64
    say("edge case", kind, "multipole", multipole)
    rng = np.random.default_rng(17)
    bits = 21
    fp64 = None
    if kind == "one":
        xyz, m = [[0.5, -0.25, 2.0]], [1.0]
    elif kind == "two":
        xyz, m = [[0, 0, 0], [1e-3, 0, 0]], [1.0, 2.0]
    elif kind == "all_equal":
        xyz, m = np.full((300, 3), 0.125), rng.uniform(0.5, 1.5, 300)
    elif kind == "pairs_equal":
        base = rng.normal(0, 1, (400, 3))
        xyz, m = np.repeat(base, 2, axis=0), rng.uniform(0.5, 1.5, 800)
    elif kind == "massless_half":
        xyz, m = rng.normal(0, 1, (1000, 3)), np.where(np.arange(1000) % 2 == 0, 0.0, 1e-3)
    elif kind == "b1_tree":
        xyz, m, bits = rng.uniform(-1, 1, (2000, 3)), np.full(2000, 5e-4), 1
    elif kind == "equal_run_3000":
        xyz = np.concatenate([np.full((3000, 3), 0.3), rng.uniform(-5, 5, (2000, 3))])
        m = rng.uniform(0.5, 1.5, 5000) * 1e-3
    elif kind == "dense_cluster":
        xyz = np.concatenate([0.37 + rng.uniform(0, 1e-3, (5000, 3)),
                              rng.uniform(-5, 5, (3000, 3))])
        m = rng.uniform(0.5, 1.5, 8000) * 1e-3
    elif kind == "equal_run_3000_bg":
        xyz = np.concatenate([np.full((3000, 3), 0.3), rng.uniform(-5, 5, (40000, 3))])
        m = rng.uniform(0.5, 1.5, 43000) * 1e-4
    elif kind.startswith("dense_cluster_bg"):
        xyz = np.concatenate([0.37 + rng.uniform(0, 1e-3, (5000, 3)),
                              rng.uniform(-5, 5, (40000, 3))])
        m = rng.uniform(0.5, 1.5, 45000) * 1e-4
        if kind.endswith("fp64"):
            fp64 = 99
    else:
        t = np.linspace(-4, 4, 1500)
        xyz, m = np.stack([t, np.zeros_like(t), np.zeros_like(t)], 1), np.full(1500, 1e-3)
    n = len(xyz)
    pos = np.zeros(n, POSITION_MASS)
    pos["position"] = np.asarray(xyz, np.float32)
    pos["mass"] = np.asarray(m, np.float32)
    vel = np.zeros(n, astro.VELOCITY_RADIUS)
    vel["velocity"] = rng.normal(0, 0.1, (n, 3)).astype(np.float32)
    col = np.zeros(n, COLOR_TYPE)
    P, Po = astro.params_physics(bits), orc.params_physics(bits)
    P.multipole = Po.multipole = multipole
    if fp64 is not None:
        P.fp64_top_layers = Po.fp64_top_layers = fp64
    ctx = small_ctx
    ctx.set_params(P)
    ctx.set_emit_octree_nodes(True)
    ctx.upload(pos, vel, col)
    ctx.build_tree()
    ctx.check()
    t = orc.build_tree(pos, Po)
    if kind.endswith("_bg") or kind.endswith("_fp64"):
        # the case reaches the leftover pass: a node of layer 11 (the last fused layer for
        # these sizes) spans more positions than a tile and its halo (2 x 1,024)
        assert n > 32768 and longest_run_without_head(t["sorted_codes"], 11) > 2048
    assert same_bytes(ctx.octree_nodes(), t["nodes"])
    if multipole == 2:  # the links the quadrupole traversal reads from the moment rows
        cap, off = ctx.layer_table()
        cnt = ctx.layer_counts()
        valid = np.concatenate([np.arange(int(off[l]), int(off[l]) + int(cnt[l]))
                                for l in range(len(cap))])
        tl = ctx.tree_links()
        q = ctx.tree_moments()[valid].astype(np.int64)
        last = tl["last"].astype(np.int64)[valid]
        last = np.where(last == 0xFFFFFFFF, 0x07FFFFFF, last)
        assert np.array_equal(q[:, 6], last | ((tl["info"].astype(np.int64)[valid] & 31) << 27))
        assert np.array_equal(q[:, 7], tl["next"].astype(np.int64)[valid])
    ctx.compute_forces()
    ctx.check()
    want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
    assert np.array_equal(ctx.forces()[:, :3], want[:, :3])
    ctx.upload(pos, vel, col)
    ctx.step(1e-3, 3)
    ctx.check()
    p2, v2, _ = ctx.download()
    p, v = pos, vel
    for _ in range(3):
        p, v, _, _ = orc.step(p, v, col, Po, 1e-3, threads=8)
    assert np.array_equal(p2["position"], p["position"])
    assert np.array_equal(v2["velocity"], v["velocity"])


@pytest.mark.parametrize("multipole", [0, 2])
@pytest.mark.parametrize("bits", [7, 21])
def test_wide_offsets_bitexact(astro, orc, small_ctx, bits, multipole):
    # the addressing of trees of 2^27 nodes or more (64-bit node offsets; with quadrupoles the
    # links from the link array and none in the moment rows), switched on for a 16k tree:
    # nodes, forces and three steps equal the oracle's bit for bit
    say("wide offsets, bits", bits, "multipole", multipole)
    ctx = small_ctx
    pos, vel, col = plummer(astro, 16384, 23)
    P, Po = astro.params_physics(bits), orc.params_physics(bits)
    P.multipole = Po.multipole = multipole
    if multipole:
        P.theta = Po.theta = 0.55
    ctx.set_params(P)
    ctx.set_emit_octree_nodes(True)
    ctx.set_wide_offsets(True)
    try:
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        t = orc.build_tree(pos, Po)
        assert same_bytes(ctx.octree_nodes(), t["nodes"])
        if multipole:  # unpacked rows (slots past each layer's count hold earlier trees)
            cap, off = ctx.layer_table()
            cnt = ctx.layer_counts()
            valid = np.concatenate([np.arange(int(off[l]), int(off[l]) + int(cnt[l]))
                                    for l in range(len(cap))])
            assert not ctx.tree_moments()[valid, 6:].any()
        want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
        assert np.array_equal(ctx.forces()[:, :3], want[:, :3])
        ctx.upload(pos, vel, col)
        ctx.step(1e-3, 3)
        ctx.check()
        p2, v2, _ = ctx.download()
    finally:
        ctx.set_wide_offsets(False)
    p, v = pos, vel
    for _ in range(3):
        p, v, _, _ = orc.step(p, v, col, Po, 1e-3, threads=8)
    assert np.array_equal(p2["position"], p["position"])
    assert np.array_equal(v2["velocity"], v["velocity"])


@pytest.mark.parametrize("top", [0, 1])
@pytest.mark.parametrize("block", [512, 1024])
@pytest.mark.parametrize("case", ["plummer_16384", "batched", "batched_quad", "whole_tree_in_lds"])
def test_traversal_blocks_and_lds_top(astro, orc, case, block, top):
    # blocks of 512 and 1,024 threads, with and without the LDS copy of the top layers: the
    # forces equal the oracle's bit for bit (one system, batched systems with their own roots --
    # a block caches the top of its middle target's system, and its targets of other systems
    # walk from device memory -- and a tree small enough that every layer, leaves included,
    # sits in LDS)
    say("traversal block", block, "top", top, case)
    if case.startswith("batched"):
        from test_gpu_batched import make_batch
        sizes = [3000, 1, 5000, 700]
        parts, pos, vel, col, off = make_batch(astro, sizes, seed0=60)
    else:
        n = 16384 if case == "plummer_16384" else 150
        pos, vel, col = plummer(astro, n, 29)
    bits = 7 if case == "whole_tree_in_lds" else 21  # 8 layers of at most 150 nodes
    P, Po = astro.params_physics(bits), orc.params_physics(bits)
    if case == "batched_quad":
        P.multipole = Po.multipole = 2
        P.theta = Po.theta = 0.55
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(P)
        if case.startswith("batched"):
            ctx.upload_systems(pos, vel, col, off)
        else:
            ctx.upload(pos, vel, col)
        ctx.set_force_block(block)
        ctx.set_force_top_cache(top)
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        f = ctx.forces()
    if case.startswith("batched"):
        for k, (p, v, c) in enumerate(parts):
            t = orc.build_tree(p, Po)
            want, _ = orc.barnes_hut(t, p, c, Po, threads=8)
            assert np.array_equal(f[off[k]:off[k + 1], :3], want[:, :3]), k
    else:
        t = orc.build_tree(pos, Po)
        want, _ = orc.barnes_hut(t, pos, col, Po, threads=8)
        assert np.array_equal(f[:, :3], want[:, :3])
