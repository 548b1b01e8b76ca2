"""astro_amd -- the Barnes-Hut step of osbo/astro on MI355X (gfx950).

The compute path is libastro.so (hand-written HIP kernels behind a C-ABI, include/astro/astro.h);
this package holds its ctypes binding and the reference-layout struct mirrors.
"""
from .structs import (AstroParams, COLOR_TYPE, OCTREE_NODE, POSITION_MASS, TREE_LINK,  # noqa
                      VELOCITY_RADIUS, INVALID_CHILD, INVALID_MORTON_CODE, SNAPSHOT_VERSION,
                      SnapshotHeader, make_particles)
from .capi import (AstroError, Context, Server, load, device_count, params_compat,  # noqa
                   params_physics, ic_plummer, ic_two_disks, ic_plummer_background,
                   ic_batch_sizes, snapshot_info, EMIT_NEVER, EMIT_EVERY_BUILD,
                   EMIT_ON_READ, REQUEST_DONE, REQUEST_FAILED, SORT_RADIX, SORT_SEEDED)
