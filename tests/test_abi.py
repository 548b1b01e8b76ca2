"""The C-ABI boundary without a GPU: libastro loads, exports every function include/astro/*.h
declares, and the Python mirrors of the shared structs have the C header's sizes and offsets."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("astro.h",):
        text = open(os.path.join(ROOT, "include", "astro", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(astro_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function(astro):
    lib = astro.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    from astro_amd.capi import exported_names
    assert set(exported_names()) == set(declared_functions())


def test_library_reports_no_device_without_gpu(astro):
    # no compute call is made here; a device count of 0 is fine in a CPU container
    assert astro.device_count() >= 0


C_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "astro/astro.h"
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
#define S(T) printf(#T " %zu\n", sizeof(T))
int main(void) {
  S(PositionMass); O(PositionMass, mass);
  S(VelocityRadius); O(VelocityRadius, radius);
  S(ColorType); O(ColorType, type);
  S(OctreeNode); O(OctreeNode, centerOfMass); O(OctreeNode, totalMass);
  O(OctreeNode, emittedColor); O(OctreeNode, emittedColorCenter); O(OctreeNode, children);
  O(OctreeNode, layer);
  S(astro_params); O(astro_params, theta); O(astro_params, type_scale); O(astro_params, kahan);
  S(astro_config); S(astro_stats);
  S(astro_snapshot_header); O(astro_snapshot_header, version); O(astro_snapshot_header, time);
  O(astro_snapshot_header, params);
  return 0;
}
"""


def test_struct_layouts_match_header(tmp_path, astro):
    from astro_amd import structs
    from astro_amd.capi import AstroConfig, AstroStats
    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    vals = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    vals = {k: int(v) for k, v in vals.items()}
    assert vals["PositionMass"] == structs.POSITION_MASS.itemsize == 32
    assert vals["PositionMass.mass"] == structs.POSITION_MASS.fields["mass"][1] == 16
    assert vals["VelocityRadius"] == structs.VELOCITY_RADIUS.itemsize
    assert vals["ColorType.type"] == structs.COLOR_TYPE.fields["type"][1] == 16
    assert vals["OctreeNode"] == structs.OCTREE_NODE.itemsize == 128
    for f in ("centerOfMass", "totalMass", "emittedColor", "emittedColorCenter", "children",
              "layer"):
        assert vals[f"OctreeNode.{f}"] == structs.OCTREE_NODE.fields[f][1], f
    assert vals["astro_params"] == ctypes.sizeof(structs.AstroParams)
    assert vals["astro_params.theta"] == structs.AstroParams.theta.offset
    assert vals["astro_params.type_scale"] == structs.AstroParams.type_scale.offset
    assert vals["astro_params.kahan"] == structs.AstroParams.kahan.offset
    assert vals["astro_config"] == ctypes.sizeof(AstroConfig)
    assert vals["astro_stats"] == ctypes.sizeof(AstroStats)
    H = structs.SnapshotHeader
    assert vals["astro_snapshot_header"] == ctypes.sizeof(H)
    for f in ("version", "time", "params"):
        assert vals[f"astro_snapshot_header.{f}"] == getattr(H, f).offset, f


def test_params_profiles_match_oracle(astro, orc):
    assert astro.params_compat().as_dict() == orc.params_compat().as_dict()
    for b in (6, 7, 21):
        assert astro.params_physics(b).as_dict() == orc.params_physics(b).as_dict()


def test_initial_conditions_are_seeded(astro):
    a = astro.ic_plummer(1000, 5)[0]
    b = astro.ic_plummer(1000, 5)[0]
    c = astro.ic_plummer(1000, 6)[0]
    assert a.tobytes() == b.tobytes() and a.tobytes() != c.tobytes()
    pos, vel, col = astro.ic_plummer(4096, 1)
    assert abs(pos["mass"].astype(np.float64).sum() - 1.0) < 1e-5
    assert np.all(np.abs(pos["position"].astype(np.float64).mean(0)) < 1e-3)
    assert np.all(col["type"] == 0)
    # virial: 2K + W ~ 0 for the Plummer model (W = -3 pi / 32 for a = M = G = 1)
    K = 0.5 * (pos["mass"][:, None] * vel["velocity"].astype(np.float64) ** 2).sum()
    assert abs(2 * K / (3 * np.pi / 32) - 1.0) < 0.1
    d = astro.ic_two_disks(1000, 2)[0]
    assert len(d) == 2000 and abs(d["mass"].astype(np.float64).sum() - 1.0) < 1e-5
    s = astro.ic_batch_sizes(256, 1024, 65536, 9)
    assert s.min() >= 1024 and s.max() <= 65536 and 5000 < s.mean() < 30000
    bg = astro.ic_plummer_background(10000, 3)[0]
    assert abs(bg["mass"].astype(np.float64).sum() - 1.0) < 1e-4
    assert np.all(np.abs(bg["position"][-1000:]) <= 10.0)


def test_create_fails_cleanly_without_gpu(astro):
    if astro.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(astro.AstroError):
        astro.Context(1024, bits_per_axis=7)


def test_snapshot_info_reads_the_documented_format(tmp_path, astro):
    # a snapshot written from Python by the format of astro.h: header, offsets, the three
    # reference-layout arrays
    from astro_amd import structs
    n = 10
    h = structs.SnapshotHeader()
    h.magic = b"ASTROSNP"
    h.version = structs.SNAPSHOT_VERSION
    h.n_particles, h.n_tree, h.n_systems, h.time = n, 7, 1, 2.5
    h.params = astro.params_physics(21)
    pos, vel, col = structs.make_particles(n)
    path = tmp_path / "s.snap"
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(np.array([0, n], np.uint32).tobytes())
        f.write(pos.tobytes() + vel.tobytes() + col.tobytes())
    info = astro.snapshot_info(path)
    assert (info["n_particles"], info["n_tree"], info["n_systems"], info["time"]) == (n, 7, 1, 2.5)
    assert info["params"].as_dict() == astro.params_physics(21).as_dict()
    bad = tmp_path / "bad.snap"
    bad.write_bytes(b"NOTASNAP" + bytes(h)[8:])
    with pytest.raises(astro.AstroError):
        astro.snapshot_info(bad)
