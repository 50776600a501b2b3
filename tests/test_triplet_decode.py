"""The triangular-index decode of the atom kernels (tri_decode,
aenet_amd/csrc/atom_common.h), restated in float32 and checked on the CPU
for every neighbor count up to 512: p -> (a, b), 0 <= a < b < n in row-major
order.  The kernel takes the square root with the hardware instruction
(1 ulp), so the check runs with the root rounded down, exactly and up by one
ulp: the two correction steps must make every decode exact either way."""
import numpy as np


def _decode(p, n, ulps):
    nn = np.float32(2 * n - 1)
    arg = (nn * nn - np.float32(8.0) * p.astype(np.float32)).astype(np.float32)
    root = np.sqrt(arg).astype(np.float32)
    if ulps:
        root = np.nextafter(root, np.float32(np.inf if ulps > 0 else -np.inf)).astype(np.float32)
    r = ((nn - root) * np.float32(0.5)).astype(np.float32).astype(np.int64)
    r = np.clip(r, 0, n - 2)
    first = lambda r: r * n - r * (r + 1) // 2  # index of row r's first pair
    r = np.where((r > 0) & (first(r) > p), r - 1, r)
    r = np.where((r < n - 2) & (first(r + 1) <= p), r + 1, r)
    return r, p - first(r) + r + 1


def test_decode_exact_for_every_count_up_to_512():
    for n in range(2, 513):
        a_ref, b_ref = np.triu_indices(n, 1)
        p = np.arange(len(a_ref), dtype=np.int64)
        for ulps in (-1, 0, 1):
            a, b = _decode(p, n, ulps)
            assert np.array_equal(a, a_ref) and np.array_equal(b, b_ref), (n, ulps)


def test_decode_catches_a_dropped_correction():
    """Without the correction steps the decode is wrong somewhere: the test
    above would see a mistake in them."""
    n = 300
    a_ref, _ = np.triu_indices(n, 1)
    p = np.arange(len(a_ref), dtype=np.int64)
    nn = np.float32(2 * n - 1)
    root = np.nextafter(np.sqrt((nn * nn - np.float32(8.0) * p.astype(np.float32)).astype(np.float32)),
                        np.float32(np.inf)).astype(np.float32)
    r = np.clip(((nn - root) * np.float32(0.5)).astype(np.float32).astype(np.int64), 0, n - 2)
    assert not np.array_equal(r, a_ref)
