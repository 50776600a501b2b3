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


def _cross_decode(p, m2, m3):
    """for_triplets' cross-type order (aenet_amd/csrc/atom_bp.hip): p =
    d m2 + a, c = (a + d) mod m3, the divisions by float32 reciprocals with
    one correction step each."""
    f32 = np.float32
    inv_m2 = f32(1.0) / f32(m2)
    inv_m3 = f32(1.0) / f32(m3)
    d = ((p.astype(f32) + f32(0.5)) * inv_m2).astype(f32).astype(np.int64)
    a = p - d * m2
    d = np.where(a < 0, d - 1, np.where(a >= m2, d + 1, d))
    a = np.where(a < 0, a + m2, np.where(a >= m2, a - m2, a))
    s = a + d
    q = ((s.astype(f32) + f32(0.5)) * inv_m3).astype(f32).astype(np.int64)
    c = s - q * m3
    c = np.where(c < 0, c + m3, np.where(c >= m3, c - m3, c))
    return a, c


def _check_cross(m2, m3):
    p = np.arange(m2 * m3, dtype=np.int64)
    a, c = _cross_decode(p, m2, m3)
    assert a.min() >= 0 and a.max() < m2 and c.min() >= 0 and c.max() < m3, (m2, m3)
    # every (a, c) exactly once
    assert len(np.unique(a * m3 + c)) == m2 * m3, (m2, m3)
    return a, c


def test_cross_type_order_is_a_bijection():
    for m2 in range(1, 97):
        for m3 in range(1, 97):
            _check_cross(m2, m3)
    for m2, m3 in ((512, 1), (1, 512), (300, 211), (512, 512), (97, 389)):
        _check_cross(m2, m3)


def test_cross_type_order_spreads_both_indexes():
    """Within 64 consecutive candidates of a 26 x 26 block (a bulk atom's
    neighbors of two species) no index appears more than 3 times, where
    row-major order repeats one index 26 times."""
    a, c = _check_cross(26, 26)
    for w in range(0, len(a) - 64, 64):
        assert np.bincount(a[w:w + 64]).max() <= 3
        assert np.bincount(c[w:w + 64]).max() <= 3
