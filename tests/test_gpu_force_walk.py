"""One walk per wave (astro_set_force_walk(ctx, 1)): the wave visits the union of its 64
targets' walks, each target taking part at its own walk's nodes only, so every force equals the
per-target walk's, and the oracle's, bit for bit.

* Plummer spheres at 21 and 7 bits per axis, targets outside the tree, the IEEE division path
  (cubes beyond +-8192) and blocks of 64 to 1,024 threads: forces equal the oracle's.
* The 500k two-disk system, 3 steps with the unreached layers left out: positions, velocities
  and forces equal the per-target walk's.
* A NaN target ends its walk at the lowest written layer (its children were left out) while the
  other 63 lanes of its wave go on: every finite target's force equals the per-target walk's.
* Where the wave walk does not apply (quadrupoles, visit counters, batches) the per-target walk
  runs: forces equal the oracle's.
Each test prints its case before touching the GPU."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


CASES = [
    ("plummer_B21_16384", 21, 16384, 0, 1.0),
    ("plummer_B7_16384", 7, 16384, 0, 1.0),
    ("plummer_B21_3001_targets_outside", 21, 3001, 2900, 1.0),
    ("plummer_B21_16384_ieee_division", 21, 16384, 0, 3.0e4),
]


@pytest.mark.parametrize("block", [64, 256, 1024])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_wave_walk_matches_oracle(astro, orc, case, block):
    name, bits, n, nt, scale = case
    say("wave walk", name, "block", block)
    pos, vel, col = astro.ic_plummer(n, 42)
    pos = pos.copy()
    pos["position"] *= np.float32(scale)
    P, Po = astro.params_physics(bits), orc.params_physics(bits)
    with astro.Context(n, bits_per_axis=21) as ctx:
        ctx.set_params(P)
        ctx.upload(pos, vel, col, n_tree=nt)
        ctx.set_force_block(block)
        ctx.set_force_walk(1)
        ctx.build_tree()
        ctx.compute_forces()
        ctx.check()
        f = ctx.forces()
    t = orc.build_tree(pos, Po, n_tree=nt or n)
    want, _ = orc.barnes_hut(t, pos, col, Po, n_tree=nt or n, threads=8)
    assert np.array_equal(f[:, :3], want[:, :3])


def steps(astro, pos, vel, col, walk, k, n_tree=0):
    with astro.Context(len(pos), bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.set_force_walk(walk)
        ctx.upload(pos, vel, col, n_tree=n_tree)
        ctx.step(1e-3, k)
        ctx.check()
        info = ctx.tree_info()
        p, v, _ = ctx.download()
        f = ctx.forces()
    return info, p, v, f


def test_two_disks_500k_wave_equals_lane(astro):
    say("wave walk: two disks 500k, 3 steps, wave vs lane")
    pos, vel, col = astro.ic_two_disks(250000, 2024)
    a = steps(astro, pos, vel, col, 1, 3)
    b = steps(astro, pos, vel, col, 0, 3)
    assert a[0][1] and a[0] == b[0]  # layers were left out in both
    assert same_bytes(a[1], b[1]) and same_bytes(a[2], b[2])
    assert np.array_equal(a[3], b[3])


def test_nan_target_ends_alone(astro):
    # the NaN target opens every node it meets, down to the lowest written layer, whose
    # children were left out: its walk ends there, the other lanes of its wave go on
    say("wave walk: 60k Plummer, a target at a NaN position")
    pos, vel, col = astro.ic_plummer(60000, 12)
    pos = pos.copy()
    pos["position"][59995] = np.float32(np.nan)
    a = steps(astro, pos, vel, col, 1, 1, n_tree=59990)
    b = steps(astro, pos, vel, col, 0, 1, n_tree=59990)
    assert a[0][1], a[0]
    fa, fb = a[3][:, :3], b[3][:, :3]
    fin = np.isfinite(fb).all(axis=1)
    assert not fin[59995] and fin.sum() == len(fin) - 1
    assert np.array_equal(np.isfinite(fa).all(axis=1), fin)
    assert np.array_equal(fa[fin], fb[fin])


@pytest.mark.parametrize("kind", ["quadrupole", "visit_counters", "batched"])
def test_wave_walk_falls_back(astro, orc, kind):
    say("wave walk set, per-target walk runs:", kind)
    P, Po = astro.params_physics(21), orc.params_physics(21)
    if kind == "batched":
        from test_gpu_batched import make_batch
        parts, pos, vel, col, off = make_batch(astro, [3000, 1, 5000, 700], seed0=60)
    else:
        pos, vel, col = astro.ic_plummer(16384, 29)
    if kind == "quadrupole":
        P.multipole = Po.multipole = 2
        P.theta = Po.theta = 0.5
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(P)
        if kind == "batched":
            ctx.upload_systems(pos, vel, col, off)
        else:
            ctx.upload(pos, vel, col)
        ctx.set_force_walk(1)
        ctx.build_tree()
        if kind == "visit_counters":
            st = ctx.traversal_stats()  # writes the forces too
        else:
            ctx.compute_forces()
        ctx.check()
        f = ctx.forces()
    if kind == "batched":
        for k, (p, v, c) in enumerate(parts):
            t = orc.build_tree(p, Po)
            want, _ = orc.barnes_hut(t, p, c, Po, threads=8)
            assert np.array_equal(f[off[k]:off[k + 1], :3], want[:, :3]), k
        return
    t = orc.build_tree(pos, Po)
    want, ost = orc.barnes_hut(t, pos, col, Po, threads=8)
    assert np.array_equal(f[:, :3], want[:, :3])
    if kind == "visit_counters":
        for k in ("visits", "leaf_terms", "monopole_terms", "opened"):
            assert st[k] == ost[k], k
