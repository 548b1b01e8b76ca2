"""Batched systems, the continuous-batching server, FP64 diagnostics and the autotuner on the
GPU.  A system's result must not depend on what it is batched with: each system has its own
cube, keys, tree and root, and no reduction crosses systems, so results are compared bit for
bit against the same system run alone and against the per-system oracle."""
import sys

import numpy as np
import pytest
from gpu_util import same_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def make_batch(astro, sizes, seed0=100):
    # np.concatenate drops the structs' padding: convert back to the 32-B reference layouts
    parts = [astro.ic_plummer(int(n), seed0 + k) for k, n in enumerate(sizes)]
    pos = np.ascontiguousarray(np.concatenate([p[0] for p in parts]), astro.POSITION_MASS)
    vel = np.ascontiguousarray(np.concatenate([p[1] for p in parts]), astro.VELOCITY_RADIUS)
    col = np.ascontiguousarray(np.concatenate([p[2] for p in parts]), astro.COLOR_TYPE)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    return parts, pos, vel, col, off


@pytest.mark.parametrize("graph,multipole", [(True, 0), (False, 0), (True, 2)])
def test_batched_equals_alone_and_oracle(astro, orc, graph, multipole):
    say("batched vs alone, graph", graph, "multipole", multipole)
    sizes = [1024, 3000, 1500, 7777, 2048, 1]
    parts, pos, vel, col, off = make_batch(astro, sizes)
    P = astro.params_physics(21)
    P.multipole = multipole
    steps = 3
    with astro.Context(len(pos), bits_per_axis=21, max_systems=8) as ctx:
        ctx.set_params(P)
        ctx.set_use_graph(graph)
        ctx.upload_systems(pos, vel, col, off)
        ctx.step(1e-3, steps)
        ctx.check()
        pb, vb, _ = ctx.download()
    Po = orc.params_physics(21)
    Po.multipole = multipole
    with astro.Context(max(sizes), bits_per_axis=21) as one:
        one.set_params(P)
        for k, (p, v, c) in enumerate(parts):
            one.upload(p, v, c)
            one.step(1e-3, steps)
            one.check()
            pa, va, _ = one.download()
            sl = slice(off[k], off[k + 1])
            assert same_bytes(pb[sl], pa), f"system {k} differs from alone"
            assert same_bytes(vb[sl], va)
            po, vo = p, v
            for _ in range(steps):
                po, vo, _, _ = orc.step(po, vo, c, Po, 1e-3, threads=8)
            assert np.array_equal(pa["position"], po["position"]), f"system {k} vs oracle"


def test_batched_256_systems_properties(astro):
    say("256 systems")
    sizes = astro.ic_batch_sizes(256, 1024, 4096, 5)
    parts, pos, vel, col, off = make_batch(astro, sizes, seed0=1000)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=256) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload_systems(pos, vel, col, off)
        ctx.build_tree()
        ctx.check()
        counts = ctx.layer_counts()
        assert counts[-1] == 256  # one root per system
        assert counts[0] <= len(pos)
        keys = ctx.read("SORTED_CODES", np.uint64, len(pos))
        perm = ctx.read("SORTED_INDICES", np.uint32, len(pos))
        for k in range(0, 256, 17):
            a, b = off[k], off[k + 1]
            assert np.all((perm[a:b] >= a) & (perm[a:b] < b))  # segments stay in place
            assert np.all(keys[a + 1:b] >= keys[a:b - 1])


def run_server(astro, P, reqs, max_particles, max_systems, budget=0, dt=1e-3, **kw):
    """reqs: list of (id, tick_of_arrival, pos, vel, col, steps).  Returns results, records."""
    srv = astro.Server(P, max_particles, max_systems, budget, **kw)
    pending = sorted(reqs, key=lambda r: r[1])
    results, records, tick = {}, [], 0
    while pending or srv.counters()["active"] or srv.counters()["queued"]:
        while pending and pending[0][1] <= tick:
            rid, _, p, v, c, steps = pending.pop(0)
            srv.submit(rid, p, v, c, steps)
        srv.tick(dt)
        for rec in srv.poll():
            records.append(rec)
            results[rec["id"]] = srv.fetch(rec["id"], rec["n"])
        tick += 1
        assert tick < 10000
    srv.close()
    return results, records


def test_server_alone_vs_mix(astro):
    say("server alone vs mix")
    P = astro.params_physics(21)
    a = astro.ic_plummer(3000, 7)
    alone, _ = run_server(astro, P, [(1, 0, *a, 6)], 100000, 16)
    others = [(10 + k, k % 4, *astro.ic_plummer(1000 + 700 * k, 50 + k), 2 + 3 * k)
              for k in range(6)]
    mix, recs = run_server(astro, P, [(1, 2, *a, 6)] + others, 100000, 16)
    assert same_bytes(alone[1][0], mix[1][0])
    assert same_bytes(alone[1][1], mix[1][1])
    assert len(recs) == 7 and all(r["t_done"] >= r["t_admit"] >= r["t_submit"] for r in recs)
    # and the plain single-system path agrees
    with astro.Context(3000, bits_per_axis=21) as ctx:
        ctx.set_params(P)
        ctx.upload(*a)
        ctx.step(1e-3, 6)
        assert same_bytes(ctx.download()[0], alone[1][0])


def test_server_evicts_a_failing_request(astro):
    # failure isolation: a request whose particles leave the server's max_extent (a system 100
    # times wider than the others, all arithmetic finite) is evicted with status FAILED after
    # its first step, while the others run to completion bit-identical to running alone
    say("server evicts a failing request")
    P = astro.params_physics(21)
    ok1, ok2 = astro.ic_plummer(3000, 71), astro.ic_plummer(2000, 72)
    wide = astro.ic_plummer(1500, 73)
    wide[0]["position"] *= np.float32(100.0)
    reqs = [(1, 0, *ok1, 6), (2, 0, *wide, 6), (3, 1, *ok2, 4)]
    mix, recs = run_server(astro, P, reqs, 100000, 8, max_extent=50.0)
    d = {r["id"]: r for r in recs}
    assert d[2]["status"] == astro.REQUEST_FAILED and 1 <= d[2]["steps_done"] < 6
    assert d[1]["status"] == d[3]["status"] == astro.REQUEST_DONE
    assert d[1]["steps_done"] == 6 and d[3]["steps_done"] == 4
    for rid, (p, v, c), steps in ((1, ok1, 6), (3, ok2, 4)):
        alone, _ = run_server(astro, P, [(rid, 0, p, v, c, steps)], 100000, 8, max_extent=50.0)
        assert same_bytes(mix[rid][0], alone[rid][0]) and same_bytes(mix[rid][1], alone[rid][1])
    # non-finite input never reaches the device: refused at submit
    bad = astro.ic_plummer(100, 74)
    bad[0]["position"][5, 1] = np.nan
    srv = astro.Server(P, 100000, 8)
    with pytest.raises(astro.AstroError) as e:
        srv.submit(9, *bad, 3)
    assert e.value.code == -1  # ASTRO_ERR_ARG
    srv.close()


def test_server_sheds_load(astro):
    # with max_wait_s set, a request whose estimated wait exceeds it is refused (ASTRO_ERR_BUSY)
    # once the server has measured its throughput; the work already admitted completes
    say("server sheds load")
    P = astro.params_physics(21)
    srv = astro.Server(P, 200000, 8, max_wait_s=0.05)
    first = astro.ic_plummer(20000, 81)
    srv.submit(1, *first, 40)
    for _ in range(8):
        srv.tick(1e-3)
    srv.sync()  # the throughput is measured at the settles
    big = astro.ic_plummer(150000, 82)
    with pytest.raises(astro.AstroError) as e:
        srv.submit(2, *big, 1000)  # far more than 0.05 s of work
    assert e.value.code == -7  # ASTRO_ERR_BUSY
    while srv.counters()["active"] or srv.counters()["queued"]:
        srv.tick(1e-3)
    recs = srv.poll()
    assert [r["id"] for r in recs] == [1] and recs[0]["status"] == astro.REQUEST_DONE
    srv.close()


def test_server_admission_respects_budget(astro):
    say("server budget")
    P = astro.params_physics(21)
    srv = astro.Server(P, 50000, 8)
    total, per = srv.device_bytes()
    srv.close()
    budget = per * 6000  # room for one 5000-particle request at a time
    r1, r2 = astro.ic_plummer(5000, 1), astro.ic_plummer(5000, 2)
    _, recs = run_server(astro, P, [(1, 0, *r1, 3), (2, 0, *r2, 3)], 50000, 8, budget)
    d = {r["id"]: r for r in recs}
    assert d[2]["t_admit"] >= d[1]["t_done"]  # the second waited for memory
    srv = astro.Server(P, 50000, 8, per * 1000)
    with pytest.raises(astro.AstroError):
        srv.submit(3, *astro.ic_plummer(5000, 3), 1)  # can never fit
    srv.close()


def test_diagnostics_fp64(astro, orc):
    say("diagnostics")
    pos, vel, col = astro.ic_plummer(8192, 9)
    with astro.Context(8192, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        d = ctx.diagnostics()[0]
        ctx.check()
    e = orc.energy_momentum(pos, vel)
    assert abs(d[0] - pos["mass"].astype(np.float64).sum()) < 1e-9
    assert abs(d[1] - e[0]) < 1e-9 * abs(e[0]) + 1e-12
    assert np.allclose(d[3:6], e[1:4], atol=1e-9)
    assert np.allclose(d[6:9] / d[0], e[4:7], atol=1e-9)
    # potential energy against the FP64 direct sum on all pairs (softened)
    x = pos["position"].astype(np.float64)
    m = pos["mass"].astype(np.float64)
    W = 0.0
    for i0 in range(0, 8192, 1024):
        dx = x[i0:i0 + 1024, None, :] - x[None, :, :]
        r = np.sqrt((dx ** 2).sum(-1) + float(np.float32(0.01)))
        w = m[i0:i0 + 1024, None] * m[None, :] / r
        idx = np.arange(i0, min(i0 + 1024, 8192))
        w[idx - i0, idx] = 0.0
        W -= 0.5 * w.sum()
    assert abs(d[2] - W) < 5e-3 * abs(W)


def test_autotune_keeps_results(astro):
    say("autotune")
    pos, vel, col = astro.ic_plummer(30000, 4)
    with astro.Context(30000, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        ctx.build_tree()
        ctx.compute_forces()
        f0 = ctx.forces()
        b, k, ms = ctx.autotune(3)
        assert b in (64, 128, 256, 512, 1024) and k in (0, 8, 32, 128) and ms > 0
        for blk, chunk in ((64, 0), (128, 8), (512, 32), (256, 1), (256, 4096), (1024, 0)):
            ctx.set_force_block(blk)
            ctx.set_force_xcd_chunk(chunk)
            ctx.compute_forces()
            assert same_bytes(ctx.forces(), f0)
        with pytest.raises(astro.AstroError):
            ctx.set_force_xcd_chunk(4097)
        ctx.check()


def test_energy_drift_small(astro):
    say("energy drift")
    pos, vel, col = astro.ic_plummer(16384, 21)
    with astro.Context(16384, bits_per_axis=21) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload(pos, vel, col)
        d0 = ctx.diagnostics()[0]
        ctx.step(1e-3, 20)
        d1 = ctx.diagnostics()[0]
    E0, E1 = d0[1] + d0[2], d1[1] + d1[2]
    assert abs(E1 - E0) < 1e-3 * abs(E0)
    assert np.all(np.abs(d1[3:6] - d0[3:6]) < 1e-4)


def test_direct_fp64_matches_oracle(astro, orc):
    # the accuracy-audit direct sum: per target over its own system's particles, in FP64
    say("FP64 direct sum")
    sizes = [3000, 5000, 1]
    parts, pos, vel, col, off = make_batch(astro, sizes, seed0=300)
    tg = np.array([0, 17, 2999, 3000, 4567, 7999, 8000], np.uint32)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(astro.params_physics(21))
        ctx.upload_systems(pos, vel, col, off)
        got = ctx.direct_forces_fp64(tg)
        with pytest.raises(astro.AstroError):
            ctx.direct_forces_fp64([len(pos)])
    eps2 = float(np.float32(0.01))
    for t, g in zip(tg, got):
        k = int(np.searchsorted(off, t, side="right") - 1)
        want = orc.direct_sum_fp64(parts[k][0], [t - off[k]], 1.0, eps2)[0]
        assert np.abs(g - want).max() <= 1e-12 * max(np.linalg.norm(want), 1e-300), (t, g, want)


def test_batched_lighting_per_system(astro, orc):
    # each system is lit by its own first particles (here 2 stars per system)
    say("batched lighting")
    sizes = [500, 1200, 3]
    parts, pos, vel, col, off = make_batch(astro, sizes, seed0=700)
    rng = np.random.default_rng(9)
    col = col.copy()
    col["color"] = rng.uniform(0, 1, (len(col), 4)).astype(np.float32)
    col["type"] = 1
    for k in range(len(sizes)):
        col["type"][off[k]:off[k] + 2] = 0
        col["type"][off[k] + 2:off[k + 1]:3] = 2
    P, Po = astro.params_physics(21), orc.params_physics(21)
    P.lighting_stars = Po.lighting_stars = 2
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(P)
        ctx.upload_systems(pos, vel, col, off)
        ctx.step(1e-3, 2)
        ctx.check()
        pb, vb, cb = ctx.download()
    for k in range(len(sizes)):
        sl = slice(off[k], off[k + 1])
        p, v, c = parts[k][0], parts[k][1], col[sl]
        for _ in range(2):
            p, v, c, _, _ = orc.step(p, v, c, Po, 1e-3, threads=8, return_col=True)
        assert np.array_equal(cb[sl]["color"], c["color"]), f"system {k}"
        assert np.array_equal(cb[sl]["type"], c["type"])
        assert np.array_equal(pb[sl]["position"], p["position"])


@pytest.mark.parametrize("batched", [False, True])
def test_snapshot_resume_is_bitexact(astro, tmp_path, batched):
    # write after 3 steps, run 4 more; a fresh context resumed from the file and run 4 steps
    # ends in the same bytes (positions, velocities, colours)
    say("snapshot resume, batched", batched)
    if batched:
        sizes = [3000, 5000]
        _, pos, vel, col, off = make_batch(astro, sizes, seed0=40)
    else:
        pos, vel, col = astro.ic_plummer(8000, 41)
    col = col.copy()
    col["type"] = 1
    col["type"][:2] = 0
    col["color"] = np.random.default_rng(4).uniform(0, 1, (len(col), 4)).astype(np.float32)
    P = astro.params_physics(21)
    P.lighting_stars = 2
    path = str(tmp_path / "state.snap")
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx:
        ctx.set_params(P)
        if batched:
            ctx.upload_systems(pos, vel, col, off)
        else:
            ctx.upload(pos, vel, col)
        ctx.step(1e-3, 3)
        ctx.snapshot_write(path, 0.003)
        ctx.step(1e-3, 4)
        ctx.check()
        a = ctx.download()
    info = astro.snapshot_info(path)
    assert info["n_particles"] == len(pos) and info["n_systems"] == (2 if batched else 1)
    with astro.Context(len(pos), bits_per_axis=21, max_systems=4) as ctx2:
        assert ctx2.snapshot_read(path) == 0.003
        assert ctx2.params().as_dict() == P.as_dict()
        ctx2.step(1e-3, 4)
        ctx2.check()
        b = ctx2.download()
    for x, y in zip(a, b):
        assert same_bytes(x, y)
