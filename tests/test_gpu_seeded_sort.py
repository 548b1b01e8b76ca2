"""The seeded sort (ASTRO_SORT_SEEDED: the key sort of a step or build started from the previous
sort's order) against numpy's stable argsort of the same keys -- the radix sort's order and the
reference's (RadixSort.swift:98-185, one-bit stable splits: ascending key, ties by index) -- bit
for bit:

* over steps of the 500k two-disk system (a previous order close to the new one: every source
  fits in its bucket's slots);
* after the positions are permuted in place (a previous order unrelated to the new keys), as they
  are (the last sort's splitters still fit) and doubled or shifted (the cube changed: splitters
  sampled again from an unrelated order, buckets past the register sort, past their slots,
  gathered from the overflow list and sorted in device memory);
* with many equal keys (sources on a few hundred points);
* with fewer tree sources than particles, and at sizes around the bucket size and the bucket
  count limit (8,028,160 sources; one more falls back to the radix sort);
and whole steps with either sort give the same bytes.  Each test prints its case before
touching the GPU."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rows(pos):
    """The device rows {x, y, z, m} of reference-layout positions."""
    return np.column_stack([pos["position"], pos["mass"]]).astype(np.float32)


def check_sorted(ctx, nt):
    codes = ctx.read("MORTON_CODES", np.uint64, nt)
    perm = ctx.read("SORTED_INDICES", np.uint32, nt)
    keys = ctx.read("SORTED_CODES", np.uint64, nt)
    want = np.argsort(codes, kind="stable").astype(np.uint32)
    assert np.array_equal(perm, want)
    assert np.array_equal(keys, codes[want])


def test_seeded_over_two_disk_steps(astro):
    # steps 2.. sort from the previous order with no overflow; every step's order is the stable
    # argsort of that step's keys
    say("seeded sort: two disks 500k, 6 steps")
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    n = len(pos)
    with astro.Context(n, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        for k in range(6):
            ctx.step(1e-3, 1)
            ctx.check()
            seeded, ovf = ctx.sort_info()
            assert seeded == (k > 0), k
            assert ovf == 0, (k, ovf)
            check_sorted(ctx, n)


@pytest.mark.parametrize("dups,change", [(False, "none"), (False, "double"), (False, "shift"),
                                         (True, "double")])
def test_seeded_after_permuting_positions(astro, dups, change):
    # the previous order is unrelated to the new keys.  Unchanged cube: the last sort's
    # splitters; doubled: the cube's scale changes, the splitters are sampled again from the
    # unrelated order and buckets overflow their slots (slow path); shifted by a quarter of the
    # cube: sampled again.  With dups, 300,000 sources sit on 400 points (ties decided by index)
    say("seeded sort: permuted positions, duplicates", dups, "cube", change)
    n = 300000
    pos, vel, col = astro.ic_plummer(n, 31)
    if dups:
        rng = np.random.default_rng(3)
        pts = pos["position"][rng.choice(n, 400, replace=False)]
        pos["position"] = pts[rng.integers(0, 400, n)]
    with astro.Context(n, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info() == (False, 0)
        check_sorted(ctx, n)
        perm = np.random.default_rng(9).permutation(n)
        p2 = pos[perm].copy()
        if change == "double":
            p2["position"] *= np.float32(2.0)
        elif change == "shift":
            p2["position"] += np.float32(0.25) * np.ptp(pos["position"], axis=0).max()
        ctx.write("POSITION_MASS", rows(p2))
        ctx.build_tree()
        ctx.check()
        seeded, ovf = ctx.sort_info()
        assert seeded
        if change == "double":
            assert ovf > 0
        check_sorted(ctx, n)
        counts = ctx.layer_counts()
    # the same positions uploaded and sorted from scratch give the same tree
    with astro.Context(n, bits_per_axis=21) as ref:
        ref.set_params(astro.params_physics(21))
        ref.set_sort_mode(astro.SORT_RADIX)
        ref.upload(p2, vel[perm], col[perm])
        ref.build_tree()
        assert list(ref.layer_counts()) == list(counts)


@pytest.mark.parametrize("n,nt", [(979, 0), (980, 0), (981, 0), (1024, 0), (1025, 0), (1961, 0),
                                  (2049, 0), (5000, 0), (70001, 0), (2000, 205), (100000, 33333)])
def test_seeded_sizes(astro, n, nt):
    # one bucket, a bucket boundary, tree sources fewer than particles; a permuted previous
    # order, then two steps
    say("seeded sort: n", n, "tree sources", nt or n)
    pos, vel, col = astro.ic_plummer(n, 100 + n)
    m = nt or n
    with astro.Context(n, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col, n_tree=nt)
        ctx.build_tree()
        check_sorted(ctx, m)
        perm = np.random.default_rng(n).permutation(n)
        ctx.write("POSITION_MASS", rows(pos[perm]))
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info()[0]
        check_sorted(ctx, m)
        for _ in range(2):
            ctx.step(1e-3, 1)
            ctx.check()
            assert ctx.sort_info()[0]
            check_sorted(ctx, m)


def test_seeded_and_radix_steps_same_bytes(astro):
    # eight graph steps of a 100k Plummer sphere with each sort: the same positions and
    # velocities, bit for bit
    say("seeded vs radix: Plummer 100k, 8 steps")
    pos, vel, col = astro.ic_plummer(100000, 8)
    out = {}
    for mode in (astro.SORT_RADIX, astro.SORT_SEEDED):
        with astro.Context(len(pos), bits_per_axis=21) as ctx:
            ctx.set_params(astro.params_physics(21))
            ctx.set_sort_mode(mode)
            ctx.upload(pos, vel, col)
            ctx.step(1e-3, 8)
            ctx.check()
            assert ctx.sort_info()[0] == (mode == astro.SORT_SEEDED)
            out[mode] = ctx.download()[:2]
    assert same_bytes(out[0][0], out[1][0])
    assert same_bytes(out[0][1], out[1][1])


@pytest.mark.parametrize("n,seeded", [(8028160, True), (8028161, False)])
def test_seeded_bucket_limit(astro, n, seeded):
    # 8,192 buckets of 980 sources at most: one source more sorts with the radix sort
    say("seeded sort: bucket limit, n", n)
    pos, vel, col = astro.ic_plummer(n, 5)
    with astro.Context(n, bits_per_axis=21, emit_octree_nodes=False) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info() == (seeded, 0)
        check_sorted(ctx, n)


# ---- batches: every system gets its own buckets; the order is (system, key, index) ----------
def check_sorted_batch(ctx, sys_of):
    codes = ctx.read("MORTON_CODES", np.uint64, len(sys_of))
    perm = ctx.read("SORTED_INDICES", np.uint32, len(sys_of))
    keys = ctx.read("SORTED_CODES", np.uint64, len(sys_of))
    want = np.lexsort((codes, sys_of)).astype(np.uint32)  # stable: ties by index
    assert np.array_equal(perm, want)
    assert np.array_equal(keys, codes[want])


def batch(astro, sizes, seed0):
    from test_gpu_batched import make_batch
    parts, pos, vel, col, off = make_batch(astro, sizes, seed0=seed0)
    sys_of = np.repeat(np.arange(len(sizes)), sizes)
    return pos, vel, col, off, sys_of


BATCH_SIZES = [1, 500, 980, 981, 3000, 20000, 1961, 70001, 2]


def test_seeded_batch_steps_and_permutations(astro):
    # the first build sorts with the radix sort and seeds the next; steps and a permutation of
    # every system's positions in place (its previous order unrelated to the new keys) sort
    # from the previous order, each system among its own buckets
    say("seeded sort: batch of", len(BATCH_SIZES), "systems", BATCH_SIZES)
    pos, vel, col, off, sys_of = batch(astro, BATCH_SIZES, 40)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=16) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload_systems(pos, vel, col, off)
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info() == (False, 0)
        check_sorted_batch(ctx, sys_of)
        for k in range(3):
            ctx.step(1e-3, 1)
            ctx.check()
            assert ctx.sort_info()[0], k
            check_sorted_batch(ctx, sys_of)
        rng = np.random.default_rng(4)
        p2, _, _ = ctx.download()
        for s in range(len(BATCH_SIZES)):
            a, b = int(off[s]), int(off[s + 1])
            p2[a:b] = p2[a:b][rng.permutation(b - a)]
        ctx.write("POSITION_MASS", rows(p2))
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info()[0]
        check_sorted_batch(ctx, sys_of)


@pytest.mark.parametrize("change", ["double_one", "shift_one"])
def test_seeded_batch_one_cube_changes(astro, change):
    # one system's cube changes (its splitters are sampled again, its buckets overflow when
    # doubled); the others keep theirs
    say("seeded sort: batch, one system's cube", change)
    sizes = [4000, 50000, 9000]
    pos, vel, col, off, sys_of = batch(astro, sizes, 70)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload_systems(pos, vel, col, off)
        ctx.build_tree()
        ctx.build_tree()
        ctx.check()
        assert ctx.sort_info() == (True, 0)
        p2 = pos.copy()
        a, b = int(off[1]), int(off[2])
        seg = p2[a:b][np.random.default_rng(2).permutation(b - a)]
        if change == "double_one":
            seg["position"] *= np.float32(2.0)
        else:
            seg["position"] += np.float32(0.3) * np.ptp(seg["position"], axis=0).max()
        p2[a:b] = seg
        ctx.write("POSITION_MASS", rows(p2))
        ctx.build_tree()
        ctx.check()
        seeded, ovf = ctx.sort_info()
        assert seeded
        if change == "double_one":
            assert ovf > 0
        check_sorted_batch(ctx, sys_of)


def test_seeded_and_radix_batch_steps_same_bytes(astro):
    # six graph steps of a batch with each sort: the same positions and velocities
    say("seeded vs radix: batch of", len(BATCH_SIZES), "systems, 6 steps")
    pos, vel, col, off, _ = batch(astro, BATCH_SIZES, 90)
    out = {}
    for mode in (astro.SORT_RADIX, astro.SORT_SEEDED):
        with astro.Context(len(pos), bits_per_axis=21, max_systems=16) as ctx:
            ctx.set_params(astro.params_physics(21))
            ctx.set_sort_mode(mode)
            ctx.upload_systems(pos, vel, col, off)
            ctx.step(1e-3, 6)
            ctx.check()
            assert ctx.sort_info()[0] == (mode == astro.SORT_SEEDED)
            out[mode] = ctx.download()[:2]
    assert same_bytes(out[0][0], out[1][0])
    assert same_bytes(out[0][1], out[1][1])
