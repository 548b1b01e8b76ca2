"""Pins the CPU oracle (oracle/oracle.c) against closed forms, textbook invariants, brute force
and special cases.  The reference ships no golden vectors (AstroTests/AstroTests.swift:13-15 is
empty), so each check here is chosen so that a plausible slip -- a swapped axis, a dropped
term, an off-by-one run end, an unstable sort -- fails it."""
import numpy as np
import pytest

from astro_amd.structs import COLOR_TYPE, INVALID_MORTON_CODE, POSITION_MASS, VELOCITY_RADIUS


def particles(xyz, m=None):
    n = len(xyz)
    pos = np.zeros(n, POSITION_MASS)
    pos["position"] = np.asarray(xyz, np.float32)
    pos["mass"] = 1.0 / n if m is None else np.asarray(m, np.float32)
    return pos


def bitloop_interleave(x, y, z):
    r = 0
    for i in range(21):
        r |= ((x >> i) & 1) << (3 * i)
        r |= ((y >> i) & 1) << (3 * i + 1)
        r |= ((z >> i) & 1) << (3 * i + 2)
    return r


# ---------------------------------------------------------------- Morton (MortonCode.metal)
def test_interleave_known_answers(orc):
    assert orc.interleave(1, 0, 0) == 1
    assert orc.interleave(0, 1, 0) == 2
    assert orc.interleave(0, 0, 1) == 4
    assert orc.interleave(2, 0, 0) == 8
    assert orc.interleave(0x1FFFFF, 0x1FFFFF, 0x1FFFFF) == 0x7FFFFFFFFFFFFFFF
    assert orc.interleave(0x1FFFFF, 0, 0) == 0x1249249249249249
    # bits above 21 are masked (MortonCode.metal:9-11)
    assert orc.interleave(0x200000 | 5, 0, 0) == orc.interleave(5, 0, 0)


def test_interleave_brute_force(orc):
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1 << 21, size=(300, 3)):
        assert orc.interleave(int(x), int(y), int(z)) == bitloop_interleave(int(x), int(y), int(z))


def test_compat_quantizer_known_answers(orc):
    P = orc.params_compat()
    # (u32)(p + 1e6) & 0x1FFFFF, interleave, keep the top 18 bits (>> 45)
    pts = [(0.0, 0.0, 0.0), (-1e6, -1e6, -1e6), (1.2e6, -3e5, 4.5e5), (-2e6, 2e6, 1.0)]
    pos = particles(pts)
    got = orc.morton_codes(pos, P)
    for (x, y, z), g in zip(pts, got):
        q = [max(0, int(np.float32(np.float32(v) + np.float32(1e6)))) for v in (x, y, z)]
        assert int(g) == bitloop_interleave(*[a & 0x1FFFFF for a in q]) >> 45
    assert int(got[1]) == 0  # -1e6 + 1e6 = 0 on every axis


def test_bbox_quantizer(orc):
    P = orc.params_physics(21)
    pts = np.array([[0.0, 0.0, 0.0], [3.0, 1.0, 2.0], [1.5, 0.5, 1.0]], np.float32)
    pos = particles(pts)
    cube = orc.bbox_cube(pos)
    assert list(cube) == [0.0, 0.0, 0.0, 4.0]  # extent 3 -> power of two 4
    codes = orc.morton_codes(pos, P, cube)
    # q = trunc((p - lo) * 2^21 / 4), clamped to 2^21 - 1
    q = lambda v: min(int(v * (1 << 21) / 4.0), (1 << 21) - 1)
    for p, c in zip(pts, codes):
        assert int(c) == bitloop_interleave(q(p[0]), q(p[1]), q(p[2]))
    # an exact power-of-two extent keeps its side; the far corner clamps to the last cell
    cube2 = orc.bbox_cube(particles([[0, 0, 0], [4, 4, 4]]))
    assert cube2[3] == 4.0
    assert int(orc.morton_codes(particles([[0, 0, 0], [4, 4, 4]]), P, cube2)[1]) == (1 << 63) - 1
    # keeping B bits = the top 3B bits of the 63-bit key
    P7 = orc.params_physics(7)
    assert np.array_equal(orc.morton_codes(pos, P7, cube), codes >> np.uint64(42))


# --------------------------------------------------------- scan and sort (RadixSort.*)
def test_exclusive_scan(orc):
    rng = np.random.default_rng(2)
    for n in (1, 2, 1023, 1024, 1025, 5000):
        a = rng.integers(0, 3, n).astype(np.uint32)
        want = np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.uint32)
        assert np.array_equal(orc.exclusive_scan(a), want)


def test_split_pass_is_stable_partition(orc):
    rng = np.random.default_rng(3)
    k = rng.integers(0, 16, 500).astype(np.uint64)
    idx = np.arange(500, dtype=np.uint32)
    ko, io = orc.split_pass(k, idx, 2)
    bit = (k >> np.uint64(2)) & np.uint64(1)
    order = np.concatenate([np.nonzero(bit == 0)[0], np.nonzero(bit == 1)[0]])
    assert np.array_equal(io, idx[order]) and np.array_equal(ko, k[order])


@pytest.mark.parametrize("case", ["random", "equal", "reversed", "invalid_tail", "few_distinct",
                                  "high_bits"])
def test_split_sort_is_stable_argsort(orc, case):
    rng = np.random.default_rng(4)
    n = 777
    if case == "random":
        k = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    elif case == "equal":
        k = np.full(n, 12345, np.uint64)
    elif case == "reversed":
        k = np.arange(n, 0, -1).astype(np.uint64)
    elif case == "invalid_tail":
        k = rng.integers(0, 1000, n).astype(np.uint64)
        k[-50:] = INVALID_MORTON_CODE
        k[10] = INVALID_MORTON_CODE
    elif case == "few_distinct":
        k = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(60)
    else:
        k = (rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63)) | rng.integers(0, 3, n).astype(np.uint64)
    ks, perm = orc.split_sort(k)
    want = np.argsort(k, kind="stable")
    assert np.array_equal(perm, want.astype(np.uint32))
    assert np.array_equal(ks, k[want])


def test_count_unique(orc):
    k = np.array([1, 1, 2, 9, 9, 9, 15, 16, INVALID_MORTON_CODE, INVALID_MORTON_CODE], np.uint64)
    cnt, flags, scan, uniq = orc.count_unique(k, False)
    assert cnt == 5 and list(uniq[:cnt]) == [0, 2, 3, 6, 7]
    assert list(flags) == [1, 0, 1, 1, 0, 0, 1, 1, 0, 0]
    # aggregate compares key >> 3: {1,1,2} | {9,9,9,15} | {16}
    cnt, flags, scan, uniq = orc.count_unique(k, True)
    assert cnt == 3 and list(uniq[:cnt]) == [0, 3, 7]


# ------------------------------------------------------------------- tree (Octree.metal)
def random_positions(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    return particles(rng.normal(0, spread, (n, 3)).astype(np.float32),
                     rng.lognormal(0, 0.5, n).astype(np.float32) / n)


def test_layer_table_matches_reference_sizes(orc):
    # Renderer.swift:373-386 with 205 tree sources: 205, 205, 205, 205, 64, 8, 1 -> 893 nodes
    cap, off = orc.layer_table(205, 6)
    assert cap == [205, 205, 205, 205, 64, 8, 1]
    assert off[-1] == 892 and off[-1] + cap[-1] == 893


@pytest.mark.parametrize("bits", [6, 7, 21])
def test_tree_topology_closed_form(orc, bits):
    pos = random_positions(3000, 5 + bits)
    P = orc.params_physics(bits)
    t = orc.build_tree(pos, P, resort_layers=True)
    assert t["status"] == 0  # every per-layer re-sort was the identity (App. A Q8)
    nodes, off, cap, cnt = t["nodes"], t["off"], t["cap"], t["count"]
    codes = t["codes"]
    for l in range(bits + 1):
        want = np.unique(codes >> np.uint64(3 * l))  # ascending distinct prefixes
        layer = nodes[off[l]:off[l] + cap[l]]
        assert cnt[l] == len(want)
        assert np.array_equal(layer["mortonCode"][:cnt[l]], want)
        assert np.all(layer["mortonCode"][cnt[l]:] == INVALID_MORTON_CODE)
        assert np.all(layer["layer"][:cnt[l]] == l)
        if l > 0:
            child_codes = nodes[off[l - 1]:off[l - 1] + cnt[l - 1]]["mortonCode"]
            for g in range(cnt[l]):
                ch = layer["children"][g]
                ch = ch[ch != 0xFFFFFFFF].astype(np.int64) - off[l - 1]
                mine = np.nonzero((child_codes >> np.uint64(3)) == want[g])[0]
                assert np.array_equal(ch, mine) and 1 <= len(ch) <= 8
    assert cnt[-1] == 1  # one root, at the last offset


@pytest.mark.parametrize("fp64_top", [0, 4, 99])
def test_tree_mass_and_com(orc, fp64_top):
    pos = random_positions(2000, 7)
    P = orc.params_physics(21)
    P.fp64_top_layers = fp64_top
    t = orc.build_tree(pos, P)
    m = pos["mass"].astype(np.float64)
    x = pos["position"].astype(np.float64)
    total = m.sum()
    nodes, off, cnt = t["nodes"], t["off"], t["count"]
    for l in range(22):
        layer = nodes[off[l]:off[l] + cnt[l]]
        assert abs(layer["totalMass"].astype(np.float64).sum() - total) < 1e-5 * total
    root = nodes[off[-1]]
    com = (m[:, None] * x).sum(0) / total
    assert np.allclose(root["centerOfMass"], com, rtol=0, atol=2e-6)
    assert abs(root["totalMass"] - total) < 2e-6 * total
    if fp64_top >= 22:  # all-FP64 sums round once: the root is the correctly rounded value
        assert root["totalMass"] == np.float32(total)
    # each leaf's COM is its particles' FP64 COM to FP32 accuracy
    sc, si = t["sorted_codes"], t["sorted_idx"]
    for g in range(0, cnt[0], 97):
        code = nodes[g]["mortonCode"]
        members = si[sc == code]
        mm = m[members]
        assert np.allclose(nodes[g]["centerOfMass"], (mm[:, None] * x[members]).sum(0) / mm.sum(),
                           rtol=1e-5, atol=1e-6)


def test_last_leaf_quirk(orc):
    # two particles share the last (largest) code: the reference ends the last run at
    # numUnique, so the last leaf loses them (Octree.metal:71); the physics profile keeps them
    xyz = [[-5e5, 0, 0], [0, 0, 0], [9e5, 9e5, 9e5], [9e5, 9e5, 9e5]]
    pos = particles(xyz, [1, 2, 3, 4])
    P = orc.params_compat()
    t = orc.build_tree(pos, P)
    leaves = t["nodes"][:t["count"][0]]
    # last run starts at sorted position 2 and ends at numUnique = 3: only mass 3 remains
    assert t["count"][0] == 3 and leaves[2]["totalMass"] == 3.0
    P.quirk_last_leaf = 0
    t = orc.build_tree(pos, P)
    assert t["nodes"][2]["totalMass"] == 7.0
    # a collision before it pushes the last run's start past numUnique: the leaf is empty
    xyz = [[-5e5, 0, 0], [-5e5, 0, 0], [0, 0, 0], [9e5, 9e5, 9e5], [9e5, 9e5, 9e5]]
    t = orc.build_tree(particles(xyz, [1, 2, 3, 4, 5]), orc.params_compat())
    assert t["count"][0] == 3 and t["nodes"][2]["totalMass"] == 0.0
    assert list(t["nodes"][2]["centerOfMass"]) == [0, 0, 0]


def test_quirk_last_leaf_single_member_run(orc):
    # without collisions numUnique == n and the quirk changes nothing
    pos = particles([[-5e5, 0, 0], [0, 0, 0], [9e5, 9e5, 9e5]], [1, 2, 3])
    a = orc.build_tree(pos, orc.params_compat())["nodes"]
    P = orc.params_compat()
    P.quirk_last_leaf = 0
    b = orc.build_tree(pos, P)["nodes"]
    assert a.tobytes() == b.tobytes()


# ------------------------------------------------------------------ force (BarnesHut.metal)
def test_theta_zero_is_direct_sum_over_leaves(orc):
    pos = random_positions(400, 9)
    col = np.zeros(400, COLOR_TYPE)
    P = orc.params_physics(21)
    P.theta = 0.0  # no internal node is ever accepted: every leaf is summed
    t = orc.build_tree(pos, P)
    f, st = orc.barnes_hut(t, pos, col, P)
    assert st["monopole_terms"] == 0 and st["leaf_terms"] == 400 * 399
    ds = orc.direct_sum_fp64(pos, np.arange(400), 1.0, float(np.float32(0.01)))
    assert np.allclose(f[:, :3], ds, rtol=2e-5, atol=1e-5 * np.abs(ds).max())


def test_two_body(orc):
    pos = particles([[0, 0, 0], [3, 0, 0]], [2.0, 5.0])
    col = np.zeros(2, COLOR_TYPE)
    P = orc.params_physics(21)
    t = orc.build_tree(pos, P)
    f, _ = orc.barnes_hut(t, pos, col, P)
    r2 = 9.0 + float(np.float32(0.01))
    assert np.isclose(f[0, 0], 5.0 * 3 / r2 ** 1.5, rtol=1e-6)
    assert np.isclose(f[1, 0], -2.0 * 3 / r2 ** 1.5, rtol=1e-6)
    assert f[0, 1] == 0 and f[0, 2] == 0


def test_far_cluster_is_one_monopole(orc):
    # a tight cluster far away is accepted at its top node: the force equals one monopole term
    rng = np.random.default_rng(11)
    cl = rng.normal(0, 1e-3, (50, 3)) + [100.0, 0, 0]
    pos = particles(np.vstack([[[0, 0, 0]], cl]), np.r_[1.0, np.full(50, 0.02)])
    col = np.zeros(51, COLOR_TYPE)
    P = orc.params_physics(21)
    t = orc.build_tree(pos, P)
    f, st = orc.barnes_hut(t, pos, col, P, targets=(0, 1))
    assert st["leaf_terms"] == 0 and st["monopole_terms"] >= 1
    c = (pos["mass"][1:, None] * pos["position"][1:]).sum(0) / pos["mass"][1:].sum()
    r2 = (c ** 2).sum() + 0.01
    assert np.allclose(f[0, :3], pos["mass"][1:].sum() * c / r2 ** 1.5, rtol=1e-4)


def test_uniform_sphere_interior_field(orc):
    # inside a uniform ball of mass M, radius R: a(r) = -G M r / R^3 (shell theorem)
    rng = np.random.default_rng(12)
    n = 20000
    v = rng.normal(size=(n, 3))
    v *= (rng.uniform(size=n) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
    pos = particles(v.astype(np.float32))
    col = np.zeros(n, COLOR_TYPE)
    P = orc.params_physics(21)
    P.theta = 0.3
    P.eps2 = 1e-6
    t = orc.build_tree(pos, P)
    f, _ = orc.barnes_hut(t, pos, col, P, targets=(0, 2000), threads=4)
    r = pos["position"][:2000].astype(np.float64)
    sel = np.linalg.norm(r, axis=1) < 0.6
    a = f[:2000, :3][sel]
    radial = (a * r[sel]).sum(1) / np.linalg.norm(r[sel], axis=1)
    ratio = radial / -np.linalg.norm(r[sel], axis=1)  # 1 for the smooth ball
    assert sel.sum() > 300 and abs(ratio.mean() - 1.0) < 0.03


def test_rms_error_falls_with_theta(orc, astro):
    pos, vel, col = astro.ic_plummer(4096, 42)
    tg = np.arange(0, 4096, 8)
    ds = orc.direct_sum_fp64(pos, tg, 1.0, float(np.float32(0.01)))
    errs = []
    for th in (0.9, 0.7, 0.5, 0.3):
        P = orc.params_physics(21)
        P.theta = th
        t = orc.build_tree(pos, P)
        f, _ = orc.barnes_hut(t, pos, col, P, threads=4)
        d = f[tg, :3] - ds
        errs.append(np.sqrt((d ** 2).sum() / (ds ** 2).sum()))
    assert errs == sorted(errs, reverse=True) and errs[-1] < 2e-3 and errs[0] < 0.05


def test_compat_self_skip_and_cutoff(orc):
    # compat cells are 2^15 units wide.  Tree sources: p0 at 0 and p1 at 2e5 (type 0); targets
    # also include dust (type 2, not in the tree): d1 at x=1 and d2 at x=10, in p0's cell.
    pos = particles([[0, 0, 0], [2e5, 0, 0], [1, 0, 0], [10, 0, 0]], [1, 1, 1, 1])
    col = np.zeros(4, COLOR_TYPE)
    col["type"] = [0, 0, 2, 2]
    P = orc.params_compat()
    t = orc.build_tree(pos, P, n_tree=2)
    f, st = orc.barnes_hut(t, pos, col, P, n_tree=2)
    g = np.float32(1e11)
    term = lambda d: float(g * np.float32(d) / np.float32(np.float32(d * d) + np.float32(0.01)) ** 1.5)
    # p0 skips its own leaf and feels p1
    assert np.isclose(f[0, 0], term(2e5), rtol=1e-5)
    # d1: p0's leaf is within r^2 <= 50 (cut off), p1 adds its term
    assert np.isclose(f[2, 0], term(2e5 - 1), rtol=1e-5)
    # d2: dust has no self-skip, so p0's leaf (r^2 = 100 > 50) pulls it: clamped to -1e5
    assert f[3, 0] == np.float32(-1e5)


def test_compat_force_clamp(orc):
    pos = particles([[0, 0, 0], [4e4, 0, 0]], [1e4, 1e4])  # two cells apart: 6.25e5 unclamped
    col = np.zeros(2, COLOR_TYPE)
    P = orc.params_compat()
    t = orc.build_tree(pos, P)
    f, _ = orc.barnes_hut(t, pos, col, P)
    assert f[0, 0] == np.float32(1e5) and f[1, 0] == np.float32(-1e5)


def test_kahan_changes_little_and_is_at_least_as_accurate(orc, astro):
    pos, vel, col = astro.ic_plummer(4096, 7)
    P = orc.params_physics(21)
    t = orc.build_tree(pos, P)
    f0, _ = orc.barnes_hut(t, pos, col, P)
    P.kahan = 1
    f1, _ = orc.barnes_hut(t, pos, col, P)
    assert np.allclose(f0, f1, rtol=1e-4, atol=1e-6)
    # against the same monopole sum in FP64 the compensated sum is closer
    assert not np.array_equal(f0, f1)


# ------------------------------------------------------------------------- integrator
def test_update_spheres_closed_form(orc):
    pos = particles([[1, 2, 3]], [1])
    vel = np.zeros(1, VELOCITY_RADIUS)
    vel["velocity"] = [[0.5, -1, 2]]
    col = np.zeros(1, COLOR_TYPE)
    f = np.array([[4, 8, -2, 0]], np.float32)
    P = orc.params_compat()
    dt = np.float32(1 / 30)
    p2, v2 = orc.update_spheres(pos, vel, f, col, dt, P)
    v = (np.float32([0.5, -1, 2]) * np.float32(0.9995)) + (np.float32([4, 8, -2]) * np.float32(5)) * dt
    assert np.array_equal(v2["velocity"][0], v.astype(np.float32))
    assert np.array_equal(p2["position"][0], np.float32([1, 2, 3]) + v.astype(np.float32) * dt)


def test_step_conserves_momentum_roughly(orc, astro):
    pos, vel, col = astro.ic_plummer(2048, 3)
    P = orc.params_physics(21)
    e0 = orc.energy_momentum(pos, vel)
    p, v = pos, vel
    for _ in range(5):
        p, v, f, _ = orc.step(p, v, col, P, 1e-3)
    e1 = orc.energy_momentum(p, v)
    assert np.all(np.abs(e1[1:4] - e0[1:4]) < 1e-4)
    assert abs(e1[0] - e0[0]) < 1e-2 * e0[0]


# ------------------------------------------------- quadrupole option (multipole = 2)
def node_moments_fp64(t, pos, layer, g):
    """Second moments of node g of a layer about its stored centre of mass, from its particles."""
    node = t["nodes"][t["off"][layer] + g]
    sc = t["sorted_codes"] >> np.uint64(3 * layer)
    members = t["sorted_idx"][sc == node["mortonCode"]]
    x = pos["position"][members].astype(np.float64)
    m = pos["mass"][members].astype(np.float64)
    d = x - node["centerOfMass"][:3].astype(np.float64)
    S = (m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
    got = np.zeros((3, 3))
    e, c = node["emittedColor"], node["emittedColorCenter"]
    got[0, 0], got[1, 1], got[2, 2] = e[0], e[1], e[2]
    got[0, 1] = got[1, 0] = e[3]
    got[0, 2] = got[2, 0] = c[0]
    got[1, 2] = got[2, 1] = c[1]
    return got, S, len(members)


@pytest.mark.parametrize("bits,fp64_top", [(21, 0), (21, 99), (5, 0)])
def test_quadrupole_moments_brute_force(orc, bits, fp64_top):
    # S = sum m (x - P)(x - P)^T over each node's particles, P the node's stored COM
    # (the parallel-axis recursion over children must reproduce it); a swapped index or a
    # dropped child term fails.
    rng = np.random.default_rng(21)
    xyz = rng.normal(0, 1, (3000, 3)) * [1.0, 0.5, 0.2]
    xyz = xyz @ np.linalg.qr(rng.normal(size=(3, 3)))[0]  # general orientation: S has xz, yz
    pos = particles(xyz.astype(np.float32), rng.uniform(0.5, 1.5, 3000))
    P = orc.params_physics(bits)
    P.multipole = 2
    P.fp64_top_layers = fp64_top
    t = orc.build_tree(pos, P)
    checked = 0
    for layer in sorted({0, 1, 2} | set(range(max(bits - 9, 0), bits + 1))):
        cnt = t["count"][layer]
        for g in np.unique(np.linspace(0, cnt - 1, 12).astype(int)):
            got, S, k = node_moments_fp64(t, pos, layer, g)
            if k == 1:  # a single particle sits at its own COM, up to the COM's rounding
                x = np.abs(t["nodes"][t["off"][layer] + g]["centerOfMass"][:3]).max()
                assert np.abs(got).max() <= 2.0 * (2e-7 * x) ** 2
                continue
            tr = np.trace(S)
            assert np.abs(got - S).max() <= 2e-5 * tr + 1e-30, (layer, g, got, S)
            checked += 1
    assert checked >= 30
    # monopole mode keeps the slots zero
    P.multipole = 0
    t0 = orc.build_tree(pos, P)
    assert not t0["nodes"]["emittedColor"].any() and not t0["nodes"]["emittedColorCenter"].any()


@pytest.mark.parametrize("axis", ["along", "across"])
def test_quadrupole_pair_closed_form(orc, axis):
    # two masses m at distance +-a from their centre, target at distance R (no softening):
    # the truncated expansion is exactly 2m/R^2 (1 + 3 a^2/R^2) along the pair's axis and
    # 2m/R^2 (1 - 3/2 a^2/R^2) across it (the latter fixes the trace term's sign).
    R, a, m = 20.0, 1.0, 0.5
    xyz = [[0, 0, 0], [R - a, 0, 0], [R + a, 0, 0]] if axis == "along" else \
          [[0, 0, 0], [-a, R, 0], [a, R, 0]]
    pos = particles(xyz, [0.0, m, m])  # a massless target: the root is exactly the pair
    col = np.zeros(3, COLOR_TYPE)
    P = orc.params_physics(21)
    P.theta, P.eps2, P.multipole = 1e30, 0.0, 2
    t = orc.build_tree(pos, P)
    f, st = orc.barnes_hut(t, pos, col, P, targets=(0, 1))
    assert st["monopole_terms"] == 1 and st["visits"] == 1
    k = 0 if axis == "along" else 1
    want = 2 * m / R ** 2 * (1 + 3 * a * a / R ** 2 if axis == "along" else 1 - 1.5 * a * a / R ** 2)
    assert abs(f[0, k] - want) <= 2e-6 * want
    assert f[0, 1 - k] == 0 and f[0, 2] == 0


def test_quadrupole_beats_monopole_on_tilted_cluster(orc):
    # an anisotropic, tilted cluster seen from a general direction: the quadrupole error is a
    # higher order in size/distance, so it must be far below the monopole error; a wrong
    # off-diagonal moment would leave an error of the monopole's order
    rng = np.random.default_rng(23)
    cl = rng.normal(0, 1, (200, 3)) * [2.0, 0.6, 0.3]
    cl = cl @ np.linalg.qr(rng.normal(size=(3, 3)))[0]
    d = np.array([0.6, -0.5, 0.62])
    xyz = np.vstack([[[0, 0, 0]], cl + 60 * d / np.linalg.norm(d)])
    pos = particles(xyz.astype(np.float32), np.r_[0.0, np.full(200, 0.005)])
    col = np.zeros(201, COLOR_TYPE)
    exact = orc.direct_sum_fp64(pos, [0], 1.0, float(np.float32(0.01)))[0]
    err = {}
    for mp in (0, 2):
        P = orc.params_physics(21)
        P.theta, P.multipole = 1e30, mp
        t = orc.build_tree(pos, P)
        f, st = orc.barnes_hut(t, pos, col, P, targets=(0, 1))
        assert st["visits"] == 1
        err[mp] = np.linalg.norm(f[0, :3] - exact) / np.linalg.norm(exact)
    assert err[0] > 1e-4 and err[2] < err[0] / 30, err


def test_quadrupole_reaches_1e3_rms_at_theta_half(orc, astro):
    pos, vel, col = astro.ic_plummer(16384, 42)
    tg = np.arange(0, 16384, 8)
    ds = orc.direct_sum_fp64(pos, tg, 1.0, float(np.float32(0.01)), threads=8)
    rms = {}
    for mp, kahan in ((0, 0), (2, 0), (2, 1)):
        P = orc.params_physics(21)
        P.theta, P.multipole, P.kahan = 0.5, mp, kahan
        t = orc.build_tree(pos, P)
        f, _ = orc.barnes_hut(t, pos, col, P, threads=8)
        d = f[tg, :3] - ds
        rms[(mp, kahan)] = np.sqrt((d ** 2).sum() / (ds ** 2).sum())
    assert rms[(2, 0)] < 1e-3 and rms[(0, 0)] > 3 * rms[(2, 0)], rms
    assert abs(rms[(2, 1)] - rms[(2, 0)]) < 0.05 * rms[(2, 0)]


# ------------------------------------------- lighting pass (BarnesHut.metal:103-167)
def test_lighting_closed_form(orc):
    # one star of colour c: a planet at distance d gets c * (4e5 / d)^2, alpha 1; dust gets the
    # same light with alpha min(max(rgb), 0.25); stars keep their colours
    pos = particles([[0, 0, 0], [8e5, 0, 0], [0, 4e6, 0], [0, 0, -2e6]], [1, 1, 1, 1])
    col = np.zeros(4, COLOR_TYPE)
    col["type"] = [0, 1, 2, 2]
    col["color"][0] = [1.0, 0.5, 0.25, 1.0]
    col["color"][1:] = 7.0  # overwritten
    out = orc.lighting_pass(pos, col, 1)  # the first num_stars particles are the light sources
    assert np.array_equal(out["color"][0], col["color"][0])
    assert np.allclose(out["color"][1], [0.25, 0.125, 0.0625, 1.0], rtol=1e-6)
    assert np.allclose(out["color"][2, :3], [0.01, 0.005, 0.0025], rtol=1e-6)
    assert np.isclose(out["color"][2, 3], 0.01, rtol=1e-6)
    assert np.allclose(out["color"][3], [0.04, 0.02, 0.01, 0.04], rtol=1e-6)


def test_lighting_uses_at_most_five_stars_and_floors_distance(orc):
    rng = np.random.default_rng(3)
    xyz = np.vstack([rng.normal(0, 1e5, (7, 3)), [[0, 0, 0]]])
    pos = particles(xyz.astype(np.float32), np.ones(8))
    col = np.zeros(8, COLOR_TYPE)
    col["type"] = [0] * 7 + [1]
    col["color"][:7] = rng.uniform(0, 1, (7, 4))
    a = orc.lighting_pass(pos, col, 7)["color"][7]
    b = orc.lighting_pass(pos, col, 5)["color"][7]
    assert np.array_equal(a, b)  # stars 6 and 7 never count
    d = np.linalg.norm(xyz[:5] - xyz[7], axis=1)
    want = (col["color"][:5, :3] * (4e5 / d)[:, None] ** 2).sum(0)
    assert np.allclose(a[:3], want, rtol=1e-5) and a[3] == 1.0
    # a planet on top of a star: distance floored at 1e-9 * 4e5
    pos2 = particles([[0, 0, 0], [0, 0, 0]], [1, 1])
    col2 = np.zeros(2, COLOR_TYPE)
    col2["type"] = [0, 1]
    col2["color"][0] = [1e-18, 0, 0, 1]
    assert np.isclose(orc.lighting_pass(pos2, col2, 1)["color"][1, 0], 1.0, rtol=1e-6)


def test_lighting_per_system(orc):
    # batched: each system is lit by its own first particles only
    pos = particles([[0, 0, 0], [4e5, 0, 0], [0, 0, 0], [0, 8e5, 0]], [1, 1, 1, 1])
    col = np.zeros(4, COLOR_TYPE)
    col["type"] = [0, 1, 0, 1]
    col["color"][0] = [1, 1, 1, 1]
    col["color"][2] = [0, 2, 0, 1]
    out = orc.lighting_pass(pos, col, 1, sys_off=[0, 2, 4])["color"]
    assert np.allclose(out[1], [1, 1, 1, 1]) and np.allclose(out[3], [0, 0.5, 0, 1])


@pytest.mark.parametrize("n,seed", [(3000, 5), (800, 9)])
def test_next_link_formula_matches_lifo_walk(orc, n, seed):
    # The traversal's `next` link: the node the reference's LIFO walk (BarnesHut.metal:22-101,
    # children pushed 0..7) visits after a node's subtree.  The library computes it per head
    # position s as off[h(s)] + #{j < s : h(j) >= h(s)} - 1 (none where h(s) is the top layer),
    # h = the highest Morton digit in which a key differs from its predecessor's.  Both forms on
    # the oracle's tree of a seeded system; a slip in h (digit index), the tie rule or the
    # "- 1" fails here.
    rng = np.random.default_rng(seed)
    pos = np.zeros(n, POSITION_MASS)
    pos["position"] = rng.normal(0.0, 1.0, (n, 3)).astype(np.float32)
    pos["position"][: n // 10] = pos["position"][0]  # a run of equal keys
    pos["mass"] = 1.0
    P = orc.params_physics(21)
    t = orc.build_tree(pos, P)
    nodes, off, cnt = t["nodes"], t["off"], t["count"]
    top = len(off) - 1
    ch = nodes["children"].astype(np.int64)
    nxt = np.full(len(nodes), -2, np.int64)
    nxt[off[top]] = -1
    for l in range(top, 0, -1):
        for p in range(off[l], off[l] + cnt[l]):
            kids = [c for c in ch[p] if c != 0xFFFFFFFF]
            for i, c in enumerate(kids):
                nxt[c] = kids[i - 1] if i > 0 else nxt[p]
    sk = t["sorted_codes"].astype(np.uint64)
    h = np.full(len(sk), -1, np.int64)
    h[0] = top
    for j in range(1, len(sk)):
        x = int(sk[j]) ^ int(sk[j - 1])
        if x:
            h[j] = min((x.bit_length() - 1) // 3, top)
    f = np.full(len(sk), -3, np.int64)
    for H in range(top):
        before = np.cumsum(h >= H) - (h >= H)
        sel = h == H
        f[sel] = off[H] + before[sel] - 1
    f[h == top] = -1
    for l in range(top + 1):
        starts = np.nonzero(h >= l)[0]
        assert len(starts) == cnt[l]
        idx = np.arange(off[l], off[l] + cnt[l])
        assert np.array_equal(nxt[idx], f[starts]), l
