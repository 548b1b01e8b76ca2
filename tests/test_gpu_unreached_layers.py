"""Builds that leave out the tree layers no walk reaches (astro_set_skip_unreached_layers, on by
default): on the 500k two-disk system at theta 0.7, eps 0.1, every node of the layers below 11
passes the opening test for any finite target, so those layers are not written in a step.

* Steps with the layers left out equal steps that write every layer, bit for bit, and the tree
  read afterwards (written again in full on read) equals the full build's, byte for byte.
* A non-finite target (a NaN or infinite position) walks the written layers only: its walk ends
  at the lowest one (whose last child is 0xFFFFFFFF), the step ends, and every finite target's
  force equals the full build's.
* Quadrupoles, batches and the compat profile write every layer.
Each test prints its case before touching the GPU."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def valid_nodes(ctx):
    """The valid nodes of every layer (slots past a layer's count are left over from earlier
    builds): TREE_POS and TREE_LINKS rows, layer by layer."""
    _, off = ctx.layer_table()
    counts = ctx.layer_counts()
    tp = ctx.read("TREE_POS", np.float32, 4 * (int(off[-1]) + int(counts[-1]))).reshape(-1, 4)
    tl = ctx.read("TREE_LINKS", np.uint32, 4 * (int(off[-1]) + int(counts[-1]))).reshape(-1, 4)
    rows = [np.arange(int(o), int(o) + int(c)) for o, c in zip(off, counts)]
    idx = np.concatenate(rows)
    return tp[idx], tl[idx]


def run(astro, pos, vel, col, skip, steps, n_tree=0):
    with astro.Context(len(pos), bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.set_skip_unreached_layers(skip)
        ctx.upload(pos, vel, col, n_tree=n_tree)
        ctx.step(1e-3, steps)
        ctx.check()
        info = ctx.tree_info()
        p, v, _ = ctx.download()
        f = ctx.forces()
        tree = valid_nodes(ctx)
        counts = list(ctx.layer_counts())
    return info, p, v, f, tree, counts


def test_two_disks_skip_equals_full(astro):
    say("unreached layers: two disks 500k, 5 steps, skip vs full")
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    a = run(astro, pos, vel, col, True, 5)
    b = run(astro, pos, vel, col, False, 5)
    low, partial = a[0]
    assert partial and low >= 8, a[0]  # layers 0 .. low-1 were left out
    assert b[0] == (0, False)
    assert same_bytes(a[1], b[1]) and same_bytes(a[2], b[2])
    assert np.array_equal(a[3], b[3])
    assert a[5] == b[5]
    # the tree read after the steps (every layer's valid nodes) is the full build's
    assert same_bytes(a[4][0], b[4][0]) and np.array_equal(a[4][1], b[4][1])


@pytest.mark.parametrize("bad", ["nan", "inf"])
def test_non_finite_target_walk_ends(astro, bad):
    # a target outside the tree sources (the last 10 of 60,000) at a NaN or infinite position
    say("unreached layers: 60k Plummer, a target at a", bad, "position")
    pos, vel, col = astro.ic_plummer(60000, 12)
    pos = pos.copy()
    pos["position"][59995] = np.float32(np.nan if bad == "nan" else np.inf)
    a = run(astro, pos, vel, col, True, 1, n_tree=59990)
    b = run(astro, pos, vel, col, False, 1, n_tree=59990)
    assert a[0][1], a[0]  # layers were left out
    fa, fb = a[3][:, :3], b[3][:, :3]
    fin = np.isfinite(fb).all(axis=1)
    assert np.array_equal(np.isfinite(fa).all(axis=1), fin)
    assert np.array_equal(fa[fin], fb[fin])


@pytest.mark.parametrize("form", ["quadrupole", "compat"])
def test_forms_that_write_every_layer(astro, form):
    say("unreached layers: every layer written for", form)
    pos, vel, col = astro.ic_plummer(50000, 3)
    if form == "quadrupole":
        P = astro.params_physics(21)
        P.multipole = 2
    else:
        P = astro.params_compat()
    with astro.Context(len(pos), bits_per_axis=21) as ctx:
        ctx.set_params(P)
        ctx.upload(pos, vel, col)
        ctx.step(1e-3, 2)
        ctx.check()
        assert ctx.tree_info() == (0, False)
