"""The traversal's links in the compact nodes, checked on the host after a build and before any
traversal reads them: every layer's child runs (last child, count) tile the layer below in
order, and every node's `next` is the node the reference's LIFO walk (BarnesHut.metal:22-101,
children pushed 0..7, so the last child is visited first) visits after the node's subtree:
its previous sibling, or for a first child its parent's next, and none (0xFFFFFFFF) after a
root.  The expected links are derived top-down from the runs alone."""
import sys

import numpy as np
import pytest

from astro_amd.structs import COLOR_TYPE, POSITION_MASS

pytestmark = pytest.mark.gpu

END = 0xFFFFFFFF


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def check_links(ctx, nsys, multipole=0):
    tn = ctx.tree_links()
    cap, off = ctx.layer_table()
    off = [int(v) for v in off]
    cnt = ctx.layer_counts()
    top = len(cap) - 1
    assert int(cnt[top]) == nsys
    info = tn["info"].astype(np.int64)
    last = tn["last"].astype(np.int64)
    exp = np.full(len(tn), -2, np.int64)
    exp[off[top]:off[top] + nsys] = END
    for l in range(top, -1, -1):
        idx = np.arange(off[l], off[l] + int(cnt[l]))
        assert np.all((info[idx] & 31) == l), l
        if l == 0:
            assert np.all(last[idx] == END)
            break
        nch = (info[idx] >> 5) & 15
        assert np.all((nch >= 1) & (nch <= 8)), l
        first = last[idx] - nch + 1
        n_below = int(cnt[l - 1])
        assert first[0] == off[l - 1], l
        assert np.array_equal(first[1:], last[idx][:-1] + 1), l
        assert last[idx][-1] == off[l - 1] + n_below - 1, l
        kids = np.arange(off[l - 1], off[l - 1] + n_below)
        par = np.repeat(idx, nch)
        is_first = kids == np.repeat(first, nch)
        exp[kids] = np.where(is_first, exp[par], kids - 1)
    valid = np.concatenate([np.arange(off[l], off[l] + int(cnt[l])) for l in range(top + 1)])
    got = tn["next"].astype(np.int64)[valid]
    bad = np.nonzero(got != exp[valid])[0]
    assert bad.size == 0, f"{bad.size} wrong next links, first at node {valid[bad[0]]}"
    if multipole == 2 and off[top] + int(cap[top]) < 1 << 27:
        # the quadrupole traversal reads the links from the second moment row instead
        q = ctx.tree_moments()[valid].astype(np.int64)
        last_v = np.where(last[valid] == END, 0x07FFFFFF, last[valid])
        assert np.array_equal(q[:, 6], last_v | ((info[valid] & 31) << 27))
        assert np.array_equal(q[:, 7], tn["next"].astype(np.int64)[valid])


def compat_system(n, nt):
    rng = np.random.default_rng(5)
    pos = np.zeros(n, POSITION_MASS)
    pos["position"] = rng.uniform(-2e6, 2e6, (n, 3)).astype(np.float32)
    pos["mass"] = np.where(np.arange(n) < 5, 1e4, np.where(np.arange(n) < nt, 1e3, 1.0))
    col = np.zeros(n, COLOR_TYPE)
    col["type"] = np.where(np.arange(n) < 5, 0, np.where(np.arange(n) < nt, 1, 2))
    return pos, col


@pytest.mark.parametrize("bits,n,multipole", [(21, 16384, 0), (21, 3001, 0), (7, 16384, 0),
                                              (6, 4000, 0), (21, 16384, 2), (21, 1, 0),
                                              (21, 2, 0)])
def test_links_one_system(astro, bits, n, multipole):
    say("links", bits, n, multipole)
    pos, vel, col = astro.ic_plummer(n, 42)
    with astro.Context(max(n, 16), bits_per_axis=bits) as ctx:
        P = astro.params_physics(bits)
        P.multipole = multipole
        ctx.set_params(P)
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.check()
        check_links(ctx, 1, multipole)


def test_links_compat_subset(astro):
    # the reference's compat parameters (its last-leaf quirk) on 205 tree sources of 2,000
    say("links compat 205 of 2000")
    pos, col = compat_system(2000, 205)
    with astro.Context(2000, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_compat())
        ctx.upload(pos, None, col, n_tree=205)
        ctx.build_tree()
        ctx.check()
        check_links(ctx, 1)


@pytest.mark.parametrize("multipole", [0, 2])
def test_links_two_disks_500k(astro, multipole):
    # with quadrupoles the links packed into the moment rows come from the fused build too
    say("links two disks 500k, multipole", multipole)
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    with astro.Context(len(pos), bits_per_axis=21, emit_octree_nodes=False) as ctx:
        P = astro.params_physics(21)
        P.multipole = multipole
        ctx.set_params(P)
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.check()
        check_links(ctx, 1, multipole)


@pytest.mark.parametrize("multipole", [0, 2])
def test_links_batched(astro, multipole):
    say("links batched", multipole)
    sizes = [1024, 3000, 1500, 7777, 2048, 1]
    parts = [astro.ic_plummer(int(n), 100 + k) for k, n in enumerate(sizes)]
    pos = np.ascontiguousarray(np.concatenate([p[0] for p in parts]), astro.POSITION_MASS)
    vel = np.ascontiguousarray(np.concatenate([p[1] for p in parts]), astro.VELOCITY_RADIUS)
    col = np.ascontiguousarray(np.concatenate([p[2] for p in parts]), astro.COLOR_TYPE)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=8) as ctx:
        P = astro.params_physics(21)
        P.multipole = multipole
        ctx.set_params(P)
        ctx.upload_systems(pos, vel, col, off)
        ctx.build_tree()
        ctx.check()
        check_links(ctx, len(sizes), multipole)
