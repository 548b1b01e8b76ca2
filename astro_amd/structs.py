"""NumPy and ctypes mirrors of include/astro/bridge.h (the reference's Astro/bridge.h:17-65).

Every 3-vector occupies a 16-byte slot, as simd_float3 does in the reference; the fourth lane
is padding.  Sizes and offsets are checked against the C header by tests/test_abi.py.
"""
import ctypes

import numpy as np

POSITION_MASS = np.dtype({"names": ["position", "mass"],
                          "formats": [("<f4", 3), "<f4"],
                          "offsets": [0, 16], "itemsize": 32})
VELOCITY_RADIUS = np.dtype({"names": ["velocity", "radius"],
                            "formats": [("<f4", 3), "<f4"],
                            "offsets": [0, 16], "itemsize": 32})
COLOR_TYPE = np.dtype({"names": ["color", "type"],
                       "formats": [("<f4", 4), "<u4"],
                       "offsets": [0, 16], "itemsize": 32})
OCTREE_NODE = np.dtype({"names": ["mortonCode", "centerOfMass", "totalMass", "emittedColor",
                                  "emittedColorCenter", "children", "layer"],
                        "formats": ["<u8", ("<f4", 3), "<f4", ("<f4", 4), ("<f4", 3),
                                    ("<u4", 8), "<u4"],
                        "offsets": [0, 16, 32, 48, 64, 80, 112], "itemsize": 128})
# links of a compact traversal node of libastro (16 B, parallel to the float4 centres and
# masses): last child, info word (layer | nchild << 5 | head level << 9), the node after the
# subtree in visiting order (0xFFFFFFFF: none), sorted position of the node's first particle
TREE_LINK = np.dtype({"names": ["last", "info", "next", "start"],
                      "formats": ["<u4", "<u4", "<u4", "<u4"],
                      "offsets": [0, 4, 8, 12], "itemsize": 16})

INVALID_MORTON_CODE = 0xFFFFFFFFFFFFFFFF
INVALID_CHILD = 0xFFFFFFFF

QUANT_OFFSET_WRAP = 0
QUANT_BBOX_POW2 = 1
SIDE_SHIFTED = 0
SIDE_EXACT = 1


class AstroParams(ctypes.Structure):
    """astro_params (include/astro/bridge.h)."""
    _fields_ = [
        ("bits_per_axis", ctypes.c_uint32),
        ("quantizer", ctypes.c_uint32),
        ("quant_offset", ctypes.c_float),
        ("side_rule", ctypes.c_uint32),
        ("theta", ctypes.c_float),
        ("eps2", ctypes.c_float),
        ("G", ctypes.c_float),
        ("r2_floor", ctypes.c_float),
        ("r_floor", ctypes.c_float),
        ("leaf_cutoff_r2", ctypes.c_float),
        ("force_clamp", ctypes.c_float),
        ("damping", ctypes.c_float),
        ("type_scale", ctypes.c_float * 3),
        ("quirk_last_leaf", ctypes.c_uint32),
        ("fp64_top_layers", ctypes.c_uint32),
        ("kahan", ctypes.c_uint32),
        ("multipole", ctypes.c_uint32),
        ("lighting_stars", ctypes.c_uint32),
    ]

    def copy(self):
        p = AstroParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self), ctypes.sizeof(self))
        return p

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "type_scale" else v
        return d


class SnapshotHeader(ctypes.Structure):
    """astro_snapshot_header (include/astro/astro.h)."""
    _fields_ = [
        ("magic", ctypes.c_char * 8),
        ("version", ctypes.c_uint32),
        ("n_particles", ctypes.c_uint32),
        ("n_tree", ctypes.c_uint32),
        ("n_systems", ctypes.c_uint32),
        ("time", ctypes.c_double),
        ("params", AstroParams),
    ]


SNAPSHOT_VERSION = 1


def make_particles(n):
    """Zeroed reference-layout particle arrays (PositionMass, VelocityRadius, ColorType)."""
    return (np.zeros(n, POSITION_MASS), np.zeros(n, VELOCITY_RADIUS), np.zeros(n, COLOR_TYPE))


def ptr(a):
    """ctypes void pointer to a contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)
