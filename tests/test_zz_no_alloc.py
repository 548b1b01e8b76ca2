"""astro_scan_noalloc (no context, no allocation) against astro_scan, bit for bit, through the
C-ABI: two calls in a row, the second on a smaller input, then the size limit."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def device_sync():
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def test_scan_noalloc_matches_scan(astro):
    lib = astro.load()
    n = 1_000_003
    rng = np.random.default_rng(31)
    with astro.Context(n, bits_per_axis=21) as ctx:
        d_in, _ = ctx.buffer("FLAGS")
        d_ref, _ = ctx.buffer("SCAN")
        d_out, _ = ctx.buffer("UNIQUE_INDICES")
        for k, length in enumerate((n, 4097)):  # the second call smaller than the first
            say("scan_noalloc call", k, "length", length)
            a = rng.integers(0, 7, length).astype(np.uint32)
            ctx.write("FLAGS", a)
            ctx.scan("FLAGS", "SCAN", length)
            ctx.sync()
            rc = lib.astro_scan_noalloc(ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), length, None)
            assert rc == 0
            device_sync()
            ctx.check()
            want = ctx.read("SCAN", np.uint32, length)
            got = ctx.read("UNIQUE_INDICES", np.uint32, length)
            assert np.array_equal(got, want)
            assert np.array_equal(want, np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.uint32))
        # beyond 16,384 tiles of 2,048 values: refused before anything runs
        assert lib.astro_scan_noalloc(ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                      16384 * 2048 + 1, None) == -3  # ASTRO_ERR_CAPACITY
        assert lib.astro_scan_noalloc(None, None, 0, None) == 0
