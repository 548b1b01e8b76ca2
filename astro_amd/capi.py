"""ctypes binding of libastro's C-ABI (include/astro/astro.h).

The library is loaded from this package directory (astro_amd/libastro.so, built in-tree by
astro_amd/csrc/Makefile).  There is no fallback: if the library is missing, every entry point
raises.  Names follow the reference's dispatch surface (Renderer.swift, RadixSort.swift).
"""
import ctypes
import os

import numpy as np

from .structs import (AstroParams, SnapshotHeader, COLOR_TYPE, OCTREE_NODE, POSITION_MASS, TREE_LINK,
                      VELOCITY_RADIUS, ptr)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libastro.so")

OK, ERR_ARG, ERR_HIP, ERR_CAPACITY, ERR_STATE, ERR_DEVICE, ERR_NOMEM, ERR_BUSY = \
    0, -1, -2, -3, -4, -5, -6, -7

BUF = dict(POSITION_MASS=0, VELOCITY_RADIUS=1, TYPE=2, FORCE=3, MORTON_CODES=4,
           UNSORTED_CODES=5, UNSORTED_INDICES=6, SORTED_CODES=7, SORTED_INDICES=8,
           UNIQUE_INDICES=9, PARENT_COUNT=10, CHILD_COUNT=11, FLAGS=12, SCAN=13,
           OCTREE_NODES=14, TREE_POS=15, LAYER_COUNTS=16, LEAF_OF_SORTED=17, CUBES=18,
           TREE_LINKS=19, TREE_MOMENTS=20)
STAGES = ["bbox_morton", "sort", "build", "force", "update"]


class AstroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libastro error {code}: {msg}")
        self.code = code


class AstroConfig(ctypes.Structure):
    _fields_ = [("max_particles", ctypes.c_uint32), ("max_tree", ctypes.c_uint32),
                ("bits_per_axis", ctypes.c_uint32), ("max_systems", ctypes.c_uint32),
                ("emit_octree_nodes", ctypes.c_uint32), ("device", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_particles", ctypes.c_uint32), ("max_systems", ctypes.c_uint32),
                ("hbm_budget", ctypes.c_uint64), ("device", ctypes.c_int32),
                ("max_extent", ctypes.c_float), ("max_wait_s", ctypes.c_double)]


REQUEST_DONE, REQUEST_FAILED = 0, 1


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("n", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("status", ctypes.c_uint32),
                ("steps_done", ctypes.c_uint32)]


class TickInfo(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_uint32), ("completed", ctypes.c_uint32),
                ("active", ctypes.c_uint32), ("queued", ctypes.c_uint32),
                ("particles", ctypes.c_uint32), ("active_after", ctypes.c_uint32),
                ("step_ms", ctypes.c_double), ("failed", ctypes.c_uint32),
                ("synced", ctypes.c_uint32)]


class AstroStats(ctypes.Structure):
    _fields_ = [("visits", ctypes.c_uint64), ("leaf_terms", ctypes.c_uint64),
                ("monopole_terms", ctypes.c_uint64), ("opened", ctypes.c_uint64)]


# (name, restype, argtypes); every exported function of astro.h
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_F = ctypes.c_float
_I = ctypes.c_int
SIGNATURES = [
    ("astro_create", _I, [ctypes.POINTER(AstroConfig), ctypes.POINTER(_P)]),
    ("astro_destroy", _I, [_P]),
    ("astro_last_error", ctypes.c_char_p, [_P]),
    ("astro_device_count", _I, [ctypes.POINTER(_I)]),
    ("astro_params_compat", None, [ctypes.POINTER(AstroParams)]),
    ("astro_params_physics", None, [ctypes.POINTER(AstroParams), _U32]),
    ("astro_set_params", _I, [_P, ctypes.POINTER(AstroParams)]),
    ("astro_get_params", _I, [_P, ctypes.POINTER(AstroParams)]),
    ("astro_set_emit_octree_nodes", _I, [_P, _I]),
    ("astro_set_use_graph", _I, [_P, _I]),
    ("astro_upload", _I, [_P, _P, _P, _P, _U32, _U32]),
    ("astro_download", _I, [_P, _P, _P, _P]),
    ("astro_download_forces", _I, [_P, _P]),
    ("astro_get_buffer", _I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_U64)]),
    ("astro_memcpy_d2h", _I, [_P, _P, _P, _U64]),
    ("astro_memcpy_h2d", _I, [_P, _P, _P, _U64]),
    ("astro_layer_table", _I, [_P, ctypes.POINTER(_U32), _P, _P]),
    ("astro_sync", _I, [_P]),
    ("astro_check", _I, [_P]),
    ("astro_generate_morton_codes", _I, [_P, _U32]),
    ("astro_fill_indices", _I, [_P, _P, _U32, _U32]),
    ("astro_clear_buffer", _I, [_P, _P, _U32, _U32]),
    ("astro_clear_octree_node_buffer", _I, [_P, _P, _U32]),
    ("astro_radix_sort", _I, [_P, _P, _P, _P, _P, _U32, _P]),
    ("astro_scan", _I, [_P, _P, _P, _U32]),
    ("astro_split_pass", _I, [_P, _P, _P, _P, _P, _U32, _U32, _P]),
    ("astro_count_unique_morton_codes", _I, [_P, _I, _U32]),
    ("astro_aggregate_leaf_nodes", _I, [_P, _U32]),
    ("astro_aggregate_internal_nodes", _I, [_P, _U32, _U32, _U32, _U32]),
    ("astro_build_internal_nodes", _I, [_P, _U32]),
    ("astro_barnes_hut", _I, [_P]),
    ("astro_lighting_pass", _I, [_P, _U32]),
    ("astro_update_spheres", _I, [_P, _F]),
    ("astro_build_tree", _I, [_P]),
    ("astro_compute_forces", _I, [_P]),
    ("astro_step", _I, [_P, _F, _U32]),
    ("astro_step_timed", _I, [_P, _F, _U32, _P]),
    ("astro_traversal_stats", _I, [_P, ctypes.POINTER(AstroStats)]),
    ("astro_upload_systems", _I, [_P, _P, _P, _P, _P, _U32]),
    ("astro_system_count", _I, [_P, ctypes.POINTER(_U32)]),
    ("astro_set_force_block", _I, [_P, _U32]),
    ("astro_set_force_xcd_chunk", _I, [_P, _U32]),
    ("astro_set_force_top_cache", _I, [_P, _I]),
    ("astro_scan_noalloc", _I, [_P, _P, _U32, _P]),
    ("astro_set_wide_offsets", _I, [_P, _I]),
    ("astro_set_sort_mode", _I, [_P, _I]),
    ("astro_set_skip_unreached_layers", _I, [_P, _I]),
    ("astro_set_force_walk", _I, [_P, _I]),
    ("astro_tree_info", _I, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_I)]),
    ("astro_sort_info", _I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_U32)]),
    ("astro_autotune", _I, [_P, _U32, ctypes.POINTER(_U32), ctypes.POINTER(_U32),
                            ctypes.POINTER(_F)]),
    ("astro_diagnostics", _I, [_P, _P]),
    ("astro_direct_forces_fp64", _I, [_P, _P, _U32, _P]),
    ("astro_particle_count", _I, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    ("astro_check_division", _I, [_P, _U64, _U64, ctypes.POINTER(_U64)]),
    ("astro_check_open_test", _I, [_P, _U64, _U64, ctypes.POINTER(_U64)]),
    ("astro_check_sqrt", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("astro_check_reciprocal", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("astro_snapshot_write", _I, [_P, ctypes.c_char_p, ctypes.c_double]),
    ("astro_snapshot_read", _I, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    ("astro_snapshot_info", _I, [ctypes.c_char_p, _P]),
    ("astro_device_bytes", _I, [_P, ctypes.POINTER(_U64)]),
    ("astro_server_create", _I, [ctypes.POINTER(ServerConfig), ctypes.POINTER(AstroParams),
                                 ctypes.POINTER(_P)]),
    ("astro_server_destroy", _I, [_P]),
    ("astro_server_last_error", ctypes.c_char_p, [_P]),
    ("astro_server_device_bytes", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("astro_server_submit", _I, [_P, _U64, _P, _P, _P, _U32, _U32]),
    ("astro_server_tick", _I, [_P, _F, ctypes.POINTER(TickInfo)]),
    ("astro_server_sync", _I, [_P, ctypes.POINTER(TickInfo)]),
    ("astro_server_poll", _I, [_P, _P, _U32, ctypes.POINTER(_U32)]),
    ("astro_server_fetch", _I, [_P, _U64, _P, _P]),
    ("astro_server_counters", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                                   ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    ("astro_ic_plummer", _I, [_U32, _U64, _P, _P, _P]),
    ("astro_ic_two_disks", _I, [_U32, _U64, _F, _F, _P, _P, _P]),
    ("astro_ic_plummer_background", _I, [_U32, _U64, _F, _F, _F, _F, _F, _P, _P, _P]),
    ("astro_ic_batch_sizes", _I, [_U32, _U32, _U32, _U64, _P]),
]

_lib = None


def load():
    """Load libastro.so from the package directory; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c "
                           f"'import __graft_entry__ as g; g.build()'` (no fallback exists)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names():
    return [s[0] for s in SIGNATURES]


def device_count():
    n = ctypes.c_int(0)
    rc = load().astro_device_count(ctypes.byref(n))
    return n.value if rc == OK else 0


def params_compat():
    p = AstroParams()
    load().astro_params_compat(ctypes.byref(p))
    return p


def params_physics(bits_per_axis=21):
    p = AstroParams()
    load().astro_params_physics(ctypes.byref(p), bits_per_axis)
    return p


def _ic_arrays(n):
    return (np.zeros(n, POSITION_MASS), np.zeros(n, VELOCITY_RADIUS), np.zeros(n, COLOR_TYPE))


def _check_ic(rc, what):
    if rc != OK:
        raise AstroError(rc, f"{what} rejected its arguments")


def ic_plummer(n, seed):
    pos, vel, col = _ic_arrays(n)
    _check_ic(load().astro_ic_plummer(n, seed, ptr(pos), ptr(vel), ptr(col)), "astro_ic_plummer")
    return pos, vel, col


def ic_two_disks(n_per_disk, seed, separation=10.0, pericentre=1.0):
    pos, vel, col = _ic_arrays(2 * n_per_disk)
    _check_ic(load().astro_ic_two_disks(n_per_disk, seed, separation, pericentre, ptr(pos),
                                        ptr(vel), ptr(col)), "astro_ic_two_disks")
    return pos, vel, col


def ic_plummer_background(n, seed, bg_fraction=0.1, cube_side=20.0, slope=2.35, m_lo=0.1,
                          m_hi=10.0):
    pos, vel, col = _ic_arrays(n)
    _check_ic(load().astro_ic_plummer_background(n, seed, bg_fraction, cube_side, slope, m_lo,
                                                 m_hi, ptr(pos), ptr(vel), ptr(col)),
              "astro_ic_plummer_background")
    return pos, vel, col


def snapshot_info(path):
    """Header of a snapshot file (host only): dict of its fields."""
    h = SnapshotHeader()
    rc = load().astro_snapshot_info(os.fsencode(path), ctypes.byref(h))
    if rc != 0:
        raise AstroError(rc, f"{path} is not a readable snapshot")
    return dict(version=h.version, n_particles=h.n_particles, n_tree=h.n_tree,
                n_systems=h.n_systems, time=h.time, params=h.params.copy())


def ic_batch_sizes(n_systems, n_min, n_max, seed):
    out = np.zeros(n_systems, np.uint32)
    _check_ic(load().astro_ic_batch_sizes(n_systems, n_min, n_max, seed, ptr(out)),
              "astro_ic_batch_sizes")
    return out


# When the OctreeNode array is written (astro.h ASTRO_EMIT_*)
EMIT_NEVER, EMIT_EVERY_BUILD, EMIT_ON_READ = 0, 1, 2
# How steps and builds sort the keys (astro.h ASTRO_SORT_*)
SORT_RADIX, SORT_SEEDED = 0, 1


def _emit_mode(m):
    if m is True:
        return EMIT_ON_READ
    if m is False or m is None:
        return EMIT_NEVER
    if m not in (EMIT_NEVER, EMIT_EVERY_BUILD, EMIT_ON_READ):
        raise ValueError(f"emission mode {m!r}")
    return int(m)


class Context:
    """One libastro context: the reference's Renderer + MetalKernelsRadixSort buffers on one
    GPU (Renderer.swift:336-391, RadixSort.swift:40-67)."""

    def __init__(self, max_particles, max_tree=0, bits_per_axis=21, max_systems=1,
                 emit_octree_nodes=EMIT_ON_READ, device=0):
        self.lib = load()
        cfg = AstroConfig(max_particles, max_tree, bits_per_axis, max_systems,
                          _emit_mode(emit_octree_nodes), device)
        h = ctypes.c_void_p()
        rc = self.lib.astro_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise AstroError(rc, "astro_create failed")
        self.h = h
        self.n = 0
        self.n_tree = 0

    # -- plumbing ------------------------------------------------------------------------
    def _chk(self, rc):
        if rc != OK:
            msg = self.lib.astro_last_error(self.h)
            raise AstroError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.astro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- parameters and state ------------------------------------------------------------
    def set_params(self, p):
        self._chk(self.lib.astro_set_params(self.h, ctypes.byref(p)))

    def params(self):
        p = AstroParams()
        self._chk(self.lib.astro_get_params(self.h, ctypes.byref(p)))
        return p

    def set_emit_octree_nodes(self, mode):
        """EMIT_NEVER, EMIT_EVERY_BUILD or EMIT_ON_READ (True: on read, False: never)."""
        self._chk(self.lib.astro_set_emit_octree_nodes(self.h, _emit_mode(mode)))

    def set_use_graph(self, on):
        self._chk(self.lib.astro_set_use_graph(self.h, 1 if on else 0))

    def upload(self, pos, vel=None, col=None, n_tree=0):
        pos = np.ascontiguousarray(pos, POSITION_MASS)
        vel = None if vel is None else np.ascontiguousarray(vel, VELOCITY_RADIUS)
        col = None if col is None else np.ascontiguousarray(col, COLOR_TYPE)
        self._chk(self.lib.astro_upload(self.h, ptr(pos), ptr(vel), ptr(col), len(pos), n_tree))
        self.n = len(pos)
        self.n_tree = n_tree if 0 < n_tree <= len(pos) else len(pos)

    def upload_systems(self, pos, vel, col, offsets):
        """Batched layout: system s owns particles [offsets[s], offsets[s+1])."""
        pos = np.ascontiguousarray(pos, POSITION_MASS)
        vel = None if vel is None else np.ascontiguousarray(vel, VELOCITY_RADIUS)
        col = None if col is None else np.ascontiguousarray(col, COLOR_TYPE)
        off = np.ascontiguousarray(offsets, np.uint32)
        self._chk(self.lib.astro_upload_systems(self.h, ptr(pos), ptr(vel), ptr(col), ptr(off),
                                                len(off) - 1))
        self.n = int(off[-1])
        self.n_tree = self.n

    def download(self):
        self._sync_counts()
        pos, vel, col = _ic_arrays(self.n)
        self._chk(self.lib.astro_download(self.h, ptr(pos), ptr(vel), ptr(col)))
        return pos, vel, col

    def forces(self):
        self._sync_counts()
        f = np.zeros((self.n, 4), np.float32)
        self._chk(self.lib.astro_download_forces(self.h, ptr(f)))
        return f

    def buffer(self, name):
        p = ctypes.c_void_p()
        b = ctypes.c_uint64()
        self._chk(self.lib.astro_get_buffer(self.h, BUF[name], ctypes.byref(p), ctypes.byref(b)))
        return p.value, b.value

    def read(self, name, dtype, count, offset=0):
        p, nbytes = self.buffer(name)
        dtype = np.dtype(dtype)
        if (offset + count) * dtype.itemsize > nbytes:
            raise ValueError(f"read of {name} beyond its {nbytes} bytes")
        out = np.zeros(count, dtype)
        if count:
            self._chk(self.lib.astro_memcpy_d2h(self.h, ptr(out),
                                                ctypes.c_void_p(p + offset * dtype.itemsize),
                                                out.nbytes))
        return out

    def write(self, name, arr, offset=0):
        p, nbytes = self.buffer(name)
        arr = np.ascontiguousarray(arr)
        if offset * arr.dtype.itemsize + arr.nbytes > nbytes:
            raise ValueError(f"write to {name} beyond its {nbytes} bytes")
        self._chk(self.lib.astro_memcpy_h2d(self.h, ctypes.c_void_p(p + offset * arr.itemsize),
                                            ptr(arr), arr.nbytes))

    def layer_table(self):
        n = ctypes.c_uint32()
        cap = np.zeros(32, np.uint64)
        off = np.zeros(32, np.uint64)
        self._chk(self.lib.astro_layer_table(self.h, ctypes.byref(n), ptr(cap), ptr(off)))
        return cap[:n.value].copy(), off[:n.value].copy()

    def sync(self):
        self._chk(self.lib.astro_sync(self.h))

    def check(self):
        self._chk(self.lib.astro_check(self.h))

    # -- stage calls (Renderer.swift / RadixSort.swift surface) ----------------------------
    def generate_morton_codes(self, count):
        self._chk(self.lib.astro_generate_morton_codes(self.h, count))

    def fill_indices(self, name, start, end):
        self._chk(self.lib.astro_fill_indices(self.h, ctypes.c_void_p(self.buffer(name)[0]),
                                              start, end))

    def clear_buffer(self, name, count, elem_bytes):
        self._chk(self.lib.astro_clear_buffer(self.h, ctypes.c_void_p(self.buffer(name)[0]),
                                              count, elem_bytes))

    def clear_octree_node_buffer(self, count):
        self._chk(self.lib.astro_clear_octree_node_buffer(
            self.h, ctypes.c_void_p(self.buffer("OCTREE_NODES")[0]), count))

    def radix_sort(self, keys_in, idx_in, keys_out, idx_out, estimated_length, count_buffer):
        """RadixSort.swift:98-104 sort(...): buffer names or raw device pointers."""
        dp = lambda x: ctypes.c_void_p(self.buffer(x)[0] if isinstance(x, str) else x)
        self._chk(self.lib.astro_radix_sort(self.h, dp(keys_in), dp(idx_in), dp(keys_out),
                                            dp(idx_out), estimated_length, dp(count_buffer)))

    def scan(self, d_in, d_out, length):
        dp = lambda x: ctypes.c_void_p(self.buffer(x)[0] if isinstance(x, str) else x)
        self._chk(self.lib.astro_scan(self.h, dp(d_in), dp(d_out), length))

    def split_pass(self, keys_in, idx_in, keys_out, idx_out, bit, estimated_length,
                   count_buffer):
        dp = lambda x: ctypes.c_void_p(self.buffer(x)[0] if isinstance(x, str) else x)
        self._chk(self.lib.astro_split_pass(self.h, dp(keys_in), dp(idx_in), dp(keys_out),
                                            dp(idx_out), bit, estimated_length,
                                            dp(count_buffer)))

    def count_unique_morton_codes(self, aggregate, count):
        self._chk(self.lib.astro_count_unique_morton_codes(self.h, 1 if aggregate else 0, count))

    def aggregate_leaf_nodes(self, count):
        self._chk(self.lib.astro_aggregate_leaf_nodes(self.h, count))

    def aggregate_internal_nodes(self, layer, input_offset, output_offset, node_count):
        self._chk(self.lib.astro_aggregate_internal_nodes(self.h, layer, input_offset,
                                                          output_offset, node_count))

    def build_internal_nodes(self, count):
        self._chk(self.lib.astro_build_internal_nodes(self.h, count))

    def barnes_hut(self):
        self._chk(self.lib.astro_barnes_hut(self.h))

    def lighting_pass(self, num_stars):
        self._chk(self.lib.astro_lighting_pass(self.h, num_stars))

    def update_spheres(self, dt):
        self._chk(self.lib.astro_update_spheres(self.h, dt))

    # -- fused path ------------------------------------------------------------------------
    def build_tree(self):
        self._chk(self.lib.astro_build_tree(self.h))

    def compute_forces(self):
        self._chk(self.lib.astro_compute_forces(self.h))

    def step(self, dt, nsteps=1):
        self._chk(self.lib.astro_step(self.h, dt, nsteps))

    def step_timed(self, dt, nsteps=1):
        ms = np.zeros(len(STAGES), np.float32)
        self._chk(self.lib.astro_step_timed(self.h, dt, nsteps, ptr(ms)))
        return dict(zip(STAGES, ms.tolist()))

    def traversal_stats(self):
        st = AstroStats()
        self._chk(self.lib.astro_traversal_stats(self.h, ctypes.byref(st)))
        return dict(visits=st.visits, leaf_terms=st.leaf_terms,
                    monopole_terms=st.monopole_terms, opened=st.opened)

    def set_force_block(self, block):
        self._chk(self.lib.astro_set_force_block(self.h, block))

    def set_force_xcd_chunk(self, chunk):
        self._chk(self.lib.astro_set_force_xcd_chunk(self.h, chunk))

    def set_force_top_cache(self, on):
        self._chk(self.lib.astro_set_force_top_cache(self.h, 1 if on else 0))

    def set_wide_offsets(self, on):
        self._chk(self.lib.astro_set_wide_offsets(self.h, 1 if on else 0))

    def set_skip_unreached_layers(self, on):
        """True (default): a build leaves out the tree layers no walk can reach (written in full
        when the tree is read)."""
        self._chk(self.lib.astro_set_skip_unreached_layers(self.h, 1 if on else 0))

    def set_force_walk(self, mode):
        """0: one walk per target (default); 1: one walk per wave (same forces)."""
        self._chk(self.lib.astro_set_force_walk(self.h, int(mode)))

    def tree_info(self):
        """(lowest layer the last build wrote, whether layers are left out until read)."""
        low, part = ctypes.c_uint32(), ctypes.c_int()
        self._chk(self.lib.astro_tree_info(self.h, ctypes.byref(low), ctypes.byref(part)))
        return low.value, bool(part.value)

    def set_sort_mode(self, mode):
        """SORT_RADIX (from scratch every build) or SORT_SEEDED (from the previous order)."""
        self._chk(self.lib.astro_set_sort_mode(self.h, int(mode)))

    def sort_info(self):
        """(seeded, overflow) of the last sort: whether it was the seeded one, and its sources
        that did not fit in their bucket's slots."""
        seeded, ovf = ctypes.c_int(), ctypes.c_uint32()
        self._chk(self.lib.astro_sort_info(self.h, ctypes.byref(seeded), ctypes.byref(ovf)))
        return bool(seeded.value), ovf.value

    def autotune(self, reps=3):
        """Picks the traversal block size and XCD chunk; returns (block, chunk, ms)."""
        b, k = ctypes.c_uint32(), ctypes.c_uint32()
        t = ctypes.c_float()
        self._chk(self.lib.astro_autotune(self.h, reps, ctypes.byref(b), ctypes.byref(k),
                                          ctypes.byref(t)))
        return b.value, k.value, t.value

    def diagnostics(self):
        """Per system: mass, kinetic, potential, px, py, pz, mx, my, mz (FP64)."""
        ns = ctypes.c_uint32()
        self._chk(self.lib.astro_system_count(self.h, ctypes.byref(ns)))
        out = np.zeros((ns.value, 9), np.float64)
        self._chk(self.lib.astro_diagnostics(self.h, ptr(out)))
        return out

    def direct_forces_fp64(self, targets):
        """FP64 direct-sum accelerations (n_targets, 3) of the given particle indices."""
        t = np.ascontiguousarray(targets, np.uint32)
        out = np.zeros((len(t), 3), np.float64)
        self._chk(self.lib.astro_direct_forces_fp64(self.h, ptr(t), len(t), ptr(out)))
        return out

    def snapshot_write(self, path, time=0.0):
        self._chk(self.lib.astro_snapshot_write(self.h, os.fsencode(path), time))

    def snapshot_read(self, path):
        """Loads a snapshot (parameters and particles); returns its stored time."""
        t = ctypes.c_double()
        self._chk(self.lib.astro_snapshot_read(self.h, os.fsencode(path), ctypes.byref(t)))
        self._sync_counts()
        return t.value

    def _sync_counts(self):
        """Array lengths from the library (what astro_download writes)."""
        n, nt = ctypes.c_uint32(), ctypes.c_uint32()
        self._chk(self.lib.astro_particle_count(self.h, ctypes.byref(n), ctypes.byref(nt)))
        self.n, self.n_tree = n.value, nt.value

    def check_division(self, n, seed=1):
        """Mismatches of the traversal's fast division against IEEE division on n pairs."""
        m = ctypes.c_uint64()
        self._chk(self.lib.astro_check_division(self.h, n, seed, ctypes.byref(m)))
        return m.value

    def check_sqrt(self):
        """(floats checked, mismatches) of the exhaustive square-root self-check."""
        n, m = ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.lib.astro_check_sqrt(self.h, ctypes.byref(n), ctypes.byref(m)))
        return n.value, m.value

    def check_reciprocal(self):
        """(floats checked, mismatches) of the exhaustive reciprocal self-check."""
        n, m = ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.lib.astro_check_reciprocal(self.h, ctypes.byref(n), ctypes.byref(m)))
        return n.value, m.value

    def check_open_test(self, n, seed=1):
        """Disagreements of the squared opening test with the square-root form on n cases."""
        m = ctypes.c_uint64()
        self._chk(self.lib.astro_check_open_test(self.h, n, seed, ctypes.byref(m)))
        return m.value

    def device_bytes(self):
        b = ctypes.c_uint64()
        self._chk(self.lib.astro_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    # -- convenience readers ---------------------------------------------------------------
    def octree_nodes(self):
        cap, off = self.layer_table()
        total = int(off[-1] + cap[-1]) if len(cap) else 0
        return self.read("OCTREE_NODES", OCTREE_NODE, total)

    def tree_pos(self):
        """Centre and mass (float32 x, y, z, m) of every compact node slot."""
        cap, off = self.layer_table()
        total = int(off[-1] + cap[-1]) if len(cap) else 0
        return self.read("TREE_POS", np.float32, 4 * total).reshape(total, 4)

    def tree_moments(self):
        """Second-moment rows of every compact node slot (multipole = 2): uint32 view (n, 8) of
        {Sxx, Syy, Szz, Sxy}, {Sxz, Syz, last | layer << 27, next} (links below 2^27 nodes)."""
        cap, off = self.layer_table()
        total = int(off[-1] + cap[-1]) if len(cap) else 0
        return self.read("TREE_MOMENTS", np.uint32, 8 * total).reshape(total, 8)

    def tree_links(self):
        """Links (last child, info, next, start) of every compact node slot."""
        cap, off = self.layer_table()
        total = int(off[-1] + cap[-1]) if len(cap) else 0
        return self.read("TREE_LINKS", TREE_LINK, total)

    def layer_counts(self):
        cap, _ = self.layer_table()
        return self.read("LAYER_COUNTS", np.uint32, len(cap))


class Server:
    """Continuous-batching server (astro_server_* in include/astro/astro.h)."""

    def __init__(self, params, max_particles, max_systems, hbm_budget=0, device=0,
                 max_extent=0.0, max_wait_s=0.0):
        self.lib = load()
        cfg = ServerConfig(max_particles, max_systems, hbm_budget, device, max_extent,
                           max_wait_s)
        h = ctypes.c_void_p()
        rc = self.lib.astro_server_create(ctypes.byref(cfg), ctypes.byref(params), ctypes.byref(h))
        if rc != OK:
            raise AstroError(rc, "astro_server_create failed")
        self.h = h

    def _chk(self, rc):
        if rc != OK:
            msg = self.lib.astro_server_last_error(self.h)
            raise AstroError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.astro_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_bytes(self):
        b, pp = ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(self.lib.astro_server_device_bytes(self.h, ctypes.byref(b), ctypes.byref(pp)))
        return b.value, pp.value

    def submit(self, rid, pos, vel, col, steps):
        pos = np.ascontiguousarray(pos, POSITION_MASS)
        vel = None if vel is None else np.ascontiguousarray(vel, VELOCITY_RADIUS)
        col = None if col is None else np.ascontiguousarray(col, COLOR_TYPE)
        self._chk(self.lib.astro_server_submit(self.h, rid, ptr(pos), ptr(vel), ptr(col),
                                               len(pos), steps))

    def tick(self, dt):
        ti = TickInfo()
        self._chk(self.lib.astro_server_tick(self.h, dt, ctypes.byref(ti)))
        return {f: getattr(ti, f) for f, _ in TickInfo._fields_}

    def sync(self):
        """Waits for the device and settles the enqueued ticks."""
        ti = TickInfo()
        self._chk(self.lib.astro_server_sync(self.h, ctypes.byref(ti)))
        return {f: getattr(ti, f) for f, _ in TickInfo._fields_}

    def poll(self, max_records=4096):
        buf = (Completion * max_records)()
        n = ctypes.c_uint32()
        self._chk(self.lib.astro_server_poll(self.h, ctypes.cast(buf, ctypes.c_void_p),
                                             max_records, ctypes.byref(n)))
        return [{f: getattr(buf[k], f) for f, _ in Completion._fields_} for k in range(n.value)]

    def fetch(self, rid, n):
        pos, vel = np.zeros(n, POSITION_MASS), np.zeros(n, VELOCITY_RADIUS)
        self._chk(self.lib.astro_server_fetch(self.h, rid, ptr(pos), ptr(vel)))
        return pos, vel

    def counters(self):
        t, ps = ctypes.c_uint64(), ctypes.c_uint64()
        q, a = ctypes.c_uint32(), ctypes.c_uint32()
        self._chk(self.lib.astro_server_counters(self.h, ctypes.byref(t), ctypes.byref(ps),
                                                 ctypes.byref(q), ctypes.byref(a)))
        return dict(ticks=t.value, particle_steps=ps.value, queued=q.value, active=a.value)
