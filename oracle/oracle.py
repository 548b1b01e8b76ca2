"""ctypes wrapper of the CPU oracle (oracle/liboracle.so, built from oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.  The product path never imports it.
"""
import ctypes
import os

import numpy as np

from astro_amd.structs import (AstroParams, COLOR_TYPE, OCTREE_NODE, POSITION_MASS,
                               VELOCITY_RADIUS, ptr)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcStats(ctypes.Structure):
    _fields_ = [("visits", ctypes.c_uint64), ("leaf_terms", ctypes.c_uint64),
                ("monopole_terms", ctypes.c_uint64), ("opened", ctypes.c_uint64),
                ("max_stack", ctypes.c_uint32)]


class OrcLayers(ctypes.Structure):
    _fields_ = [("layers", ctypes.c_uint32), ("cap", ctypes.c_uint64 * 32),
                ("off", ctypes.c_uint64 * 32), ("total", ctypes.c_uint64),
                ("count", ctypes.c_uint32 * 32)]


_P = ctypes.c_void_p
_U32, _U64, _F, _I, _D = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int, ctypes.c_double
_PP = ctypes.POINTER(AstroParams)
SIGS = [
    ("orc_params_compat", None, [_PP]),
    ("orc_params_physics", None, [_PP, _U32]),
    ("orc_interleave", _U64, [_U32, _U32, _U32]),
    ("orc_bbox_cube", None, [_P, _U32, _P]),
    ("orc_morton_codes", None, [_P, _U32, _PP, _P, _P]),
    ("orc_exclusive_scan", None, [_P, _P, _U32]),
    ("orc_split_pass", None, [_P, _P, _P, _P, _U32, _U32]),
    ("orc_split_sort", None, [_P, _P, _U32]),
    ("orc_count_unique", _U32, [_P, _U32, _I, _P, _P, _P]),
    ("orc_layer_table", None, [_U32, _U32, ctypes.POINTER(OrcLayers)]),
    ("orc_build_tree", _I, [_P, _U32, _PP, _P, _I, _P, ctypes.POINTER(OrcLayers), _P, _P, _P]),
    ("orc_barnes_hut", None, [_P, _U64, _P, _P, _P, _U32, _U32, _PP, _F, _U32, _U32, _P,
                              ctypes.POINTER(OrcStats), _I]),
    ("orc_update_spheres", None, [_P, _P, _P, _P, _U32, _F, _PP]),
    ("orc_lighting_pass", None, [_P, _P, _U32, _U32, _P, _U32]),
    ("orc_step", _I, [_P, _P, _P, _U32, _U32, _PP, _F, _P, ctypes.POINTER(OrcStats), _I, _P,
                      ctypes.POINTER(OrcLayers)]),
    ("orc_direct_sum_fp64", None, [_P, _U32, _P, _U32, _D, _D, _P, _P, _I]),
    ("orc_energy_momentum", None, [_P, _P, _U32, _P]),
]

_lib = None


def _pm(a, copy=False):
    """PositionMass array in the padded C layout (numpy drops padding when concatenating)."""
    return np.array(a, POSITION_MASS) if copy else np.ascontiguousarray(a, POSITION_MASS)


def _vr(a, copy=False):
    return np.array(a, VELOCITY_RADIUS) if copy else np.ascontiguousarray(a, VELOCITY_RADIUS)


def _ct(a, copy=False):
    if a is None:
        return None
    return np.array(a, COLOR_TYPE) if copy else np.ascontiguousarray(a, COLOR_TYPE)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def params_compat():
    p = AstroParams()
    lib().orc_params_compat(ctypes.byref(p))
    return p


def params_physics(bits=21):
    p = AstroParams()
    lib().orc_params_physics(ctypes.byref(p), bits)
    return p


def interleave(x, y, z):
    return lib().orc_interleave(x, y, z)


def bbox_cube(pos, n=None):
    pos = _pm(pos)
    n = len(pos) if n is None else n
    cube = np.zeros(4, np.float32)
    lib().orc_bbox_cube(ptr(pos), n, ptr(cube))
    return cube


def morton_codes(pos, params, cube=None, n=None):
    pos = _pm(pos)
    n = len(pos) if n is None else n
    if cube is None:
        cube = bbox_cube(pos, n) if params.quantizer == 1 else np.array([0, 0, 0, 1], np.float32)
    out = np.zeros(n, np.uint64)
    lib().orc_morton_codes(ptr(pos), n, ctypes.byref(params), ptr(np.asarray(cube, np.float32)),
                           ptr(out))
    return out


def exclusive_scan(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.zeros_like(a)
    lib().orc_exclusive_scan(ptr(a), ptr(out), len(a))
    return out


def split_pass(keys, idx, bit):
    keys = np.ascontiguousarray(keys, np.uint64)
    idx = np.ascontiguousarray(idx, np.uint32)
    ko, io = np.zeros_like(keys), np.zeros_like(idx)
    lib().orc_split_pass(ptr(keys), ptr(idx), ptr(ko), ptr(io), bit, len(keys))
    return ko, io


def split_sort(keys, idx=None):
    keys = np.array(keys, np.uint64)
    idx = np.arange(len(keys), dtype=np.uint32) if idx is None else np.array(idx, np.uint32)
    lib().orc_split_sort(ptr(keys), ptr(idx), len(keys))
    return keys, idx


def count_unique(sorted_keys, aggregate):
    k = np.ascontiguousarray(sorted_keys, np.uint64)
    n = len(k)
    flags, scan, uniq = (np.zeros(n, np.uint32) for _ in range(3))
    cnt = lib().orc_count_unique(ptr(k), n, 1 if aggregate else 0, ptr(flags), ptr(scan),
                                 ptr(uniq))
    return cnt, flags, scan, uniq


def layer_table(nt, bits):
    L = OrcLayers()
    lib().orc_layer_table(nt, bits, ctypes.byref(L))
    return ([int(L.cap[i]) for i in range(L.layers)], [int(L.off[i]) for i in range(L.layers)])


def build_tree(pos, params, n_tree=None, cube=None, resort_layers=False):
    """Returns dict(nodes, cap, off, count, codes, sorted_codes, sorted_idx, cube, status)."""
    pos = _pm(pos)
    nt = len(pos) if n_tree is None else n_tree
    if cube is None:
        cube = bbox_cube(pos, nt) if params.quantizer == 1 else np.array([0, 0, 0, 1], np.float32)
    cube = np.asarray(cube, np.float32)
    L = OrcLayers()
    lib().orc_layer_table(nt, params.bits_per_axis, ctypes.byref(L))
    nodes = np.zeros(max(int(L.total), 1), OCTREE_NODE)
    codes = np.zeros(max(nt, 1), np.uint64)
    sc = np.zeros(max(nt, 1), np.uint64)
    si = np.zeros(max(nt, 1), np.uint32)
    st = lib().orc_build_tree(ptr(pos), nt, ctypes.byref(params), ptr(cube),
                              1 if resort_layers else 0, ptr(nodes), ctypes.byref(L), ptr(codes),
                              ptr(sc), ptr(si))
    nl = L.layers
    return dict(nodes=nodes[:int(L.total)], cap=[int(L.cap[i]) for i in range(nl)],
                off=[int(L.off[i]) for i in range(nl)], count=[int(L.count[i]) for i in range(nl)],
                codes=codes[:nt], sorted_codes=sc[:nt], sorted_idx=si[:nt], cube=cube,
                status=st)


def barnes_hut(tree, pos, col, params, n_tree=None, targets=None, threads=0):
    """Forces (n, 4) for targets [t0, t1) (default: all); returns (force4, stats dict)."""
    pos, col = _pm(pos), _ct(col)
    ns = len(pos)
    nt = ns if n_tree is None else n_tree
    t0, t1 = (0, ns) if targets is None else targets
    f = np.zeros((ns, 4), np.float32)
    st = OrcStats()
    root = tree["off"][-1]
    lib().orc_barnes_hut(ptr(tree["nodes"]), root, ptr(pos), ptr(col), ptr(tree["codes"]), ns, nt,
                         ctypes.byref(params), float(tree["cube"][3]), t0, t1, ptr(f),
                         ctypes.byref(st), threads)
    return f, dict(visits=st.visits, leaf_terms=st.leaf_terms, monopole_terms=st.monopole_terms,
                   opened=st.opened, max_stack=st.max_stack)


def update_spheres(pos, vel, force4, col, dt, params):
    pos, vel, col = _pm(pos, True), _vr(vel, True), _ct(col)
    lib().orc_update_spheres(ptr(pos), ptr(vel), ptr(np.ascontiguousarray(force4, np.float32)),
                             ptr(col), len(pos), dt, ctypes.byref(params))
    return pos, vel


def lighting_pass(pos, col, num_stars, sys_off=None):
    """BarnesHut.metal:103-167 on a copy of col; returns the lit colours."""
    pos, col = _pm(pos), _ct(col, True)
    off = None if sys_off is None else np.ascontiguousarray(sys_off, np.uint32)
    lib().orc_lighting_pass(ptr(pos), ptr(col), len(pos), num_stars, ptr(off),
                            0 if off is None else len(off) - 1)
    return col


def step(pos, vel, col, params, dt, n_tree=None, threads=0, return_col=False):
    """One reference step on copies; returns (pos, vel, force4, stats), or
    (pos, vel, col, force4, stats) with return_col (the lighting pass updates col)."""
    pos, vel, col = _pm(pos, True), _vr(vel, True), _ct(col, True)
    ns = len(pos)
    nt = ns if n_tree is None else n_tree
    f = np.zeros((ns, 4), np.float32)
    st = OrcStats()
    rc = lib().orc_step(ptr(pos), ptr(vel), ptr(col), ns, nt, ctypes.byref(params), dt, ptr(f),
                        ctypes.byref(st), threads, None, None)
    if rc not in (0,):
        raise RuntimeError(f"orc_step returned {rc}")
    stats = dict(visits=st.visits, leaf_terms=st.leaf_terms, monopole_terms=st.monopole_terms,
                 opened=st.opened)
    if return_col:
        return pos, vel, col, f, stats
    return pos, vel, f, stats


def direct_sum_fp64(pos, targets, G=1.0, eps2=0.01, skip_keys=None, threads=0):
    pos = _pm(pos)
    t = np.ascontiguousarray(targets, np.uint32)
    out = np.zeros((len(t), 3), np.float64)
    sk = None if skip_keys is None else np.ascontiguousarray(skip_keys, np.uint64)
    lib().orc_direct_sum_fp64(ptr(pos), len(pos), ptr(t), len(t), G, eps2, ptr(sk), ptr(out),
                              threads)
    return out


def energy_momentum(pos, vel):
    pos, vel = _pm(pos), _vr(vel)
    out = np.zeros(7, np.float64)
    lib().orc_energy_momentum(ptr(pos), ptr(vel), len(pos), ptr(out))
    return out
