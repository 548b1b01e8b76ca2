"""Helpers shared by the GPU tests."""
import numpy as np


def same_bytes(a, b, chunk=1 << 24):
    """Byte equality of two arrays, compared a chunk at a time.  Returns a bool, so that a
    failing assert reports a short message rather than a diff of megabytes."""
    a = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    b = np.ascontiguousarray(b).view(np.uint8).reshape(-1)
    if a.size != b.size:
        return False
    return all(np.array_equal(a[i:i + chunk], b[i:i + chunk]) for i in range(0, a.size, chunk))
