"""Exactness of the integer shortcuts the HIP kernels take, checked on the CPU.

- search_kernel.hip mod_magic: a mod d for a < 2^31 and 512 <= d < 2^24 as
  a - q*d with q = mulhi(a, m) >> (l - 1), l = ceil(log2 d),
  m = ceil(2^(31+l) / d) (the kernel takes m as the ceiling of an f64
  quotient, and q*d as a 24x24-bit product);
- partition_kernels.hip k_hash_route: the level-1 bin p / G and p mod G for
  G < 2^16, p < 2^21 partitions and p / G at most 1024 bins, from an f32
  quotient estimate (numpy float32 rounds like v_cvt_f32_u32 and v_mul_f32,
  to nearest even; v_cvt_u32_f32 truncates like astype) and one correction
  step.

Both are compared with exact integer division, near every quotient boundary
and on random points.
"""
import numpy as np

U32 = np.uint32


def magic_constants(d: int) -> tuple[int, int]:
    """(m, shift) as the kernel derives them: l = 32 - clz(d - 1), m the
    ceiling of the f64 quotient 2^(31+l) / d."""
    l = int(d - 1).bit_length()
    m = int(np.ceil(np.ldexp(np.float64(1.0), 31 + l) / np.float64(d)))
    return m, l - 1


def mod_magic(a: np.ndarray, d: int, m: int | None = None, shift: int | None = None) -> np.ndarray:
    if m is None:
        m, shift = magic_constants(d)
    q = ((a.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)) >> np.uint64(shift)
    return (a.astype(np.uint64) - (q & 0xFFFFFF) * (d & 0xFFFFFF)) & 0xFFFFFFFF


def boundary_points(d: int, rng) -> np.ndarray:
    k = np.arange(0, (1 << 31) // d + 1, max(1, ((1 << 31) // d) // 200000), dtype=np.int64)
    base = k * d
    a = np.concatenate([base + o for o in (-2, -1, 0, 1, 2, d - 1)] +
                       [rng.integers(0, 1 << 31, 300000, dtype=np.int64), [(1 << 31) - 1]])
    return a[(a >= 0) & (a < (1 << 31))]


def test_magic_constant_from_f64_is_the_exact_ceiling():
    # every table size up to 2^16, and random ones up to 2^24
    rng = np.random.default_rng(10)
    ds = np.concatenate([np.arange(512, 1 << 16), rng.integers(1 << 16, 1 << 24, 20000)])
    for d in ds.tolist():
        l = int(d - 1).bit_length()
        m, _ = magic_constants(d)
        assert m == -((-(1 << (31 + l))) // d), d  # exact integer ceiling
        assert m < (1 << 32), d


def test_mod_magic_exact_near_boundaries_and_random():
    rng = np.random.default_rng(11)
    sizes = [512, 513, 1000, 1024, 1025, 2047, 2048, 2049, 2069, 2112, 3072, 3103, 4095, 4096, 4097, 6000, 8191,
             8192, 65537, 1 << 20, (1 << 24) - 1]
    for d in sizes:
        a = boundary_points(d, rng)
        assert np.array_equal(mod_magic(a, d), (a % d).astype(np.uint64)), d


def test_mod_magic_exhaustive_small_range():
    # every a below 2^22 for a few table sizes around P = 2048
    a = np.arange(0, 1 << 22, dtype=np.int64)
    for d in (512, 2048, 2069, 3072):
        assert np.array_equal(mod_magic(a, d), (a % d).astype(np.uint64)), d


def test_mod_magic_check_catches_a_weaker_constant():
    # the same check fails for m = floor(2^(31+l) / d) (one too small), and
    # for one bit less of shift: it is sharp enough to catch either slip
    rng = np.random.default_rng(13)
    for d in (2069, 3103, 6000):
        a = boundary_points(d, rng)
        m, shift = magic_constants(d)
        assert not np.array_equal(mod_magic(a, d, m - 1, shift), (a % d).astype(np.uint64)), d
        assert not np.array_equal(mod_magic(a, d, m, shift - 1), (a % d).astype(np.uint64)), d


def bin_f32(p: np.ndarray, G: int):
    rG = np.float32(1.0) / np.float32(G)
    c = (p.astype(np.float32) * rG).astype(np.int64)
    r = p.astype(np.int64) - (c & 0xFFFFFF) * (G & 0xFFFFFF)
    lo, hi = r < 0, r >= G
    c = c - lo + hi
    r = r + lo * G - hi * G
    return c, r


def test_level1_bin_exact():
    rng = np.random.default_rng(12)
    Gs = np.unique(np.concatenate([np.arange(1, 300), rng.integers(300, 1 << 16, 700), [15376, 65535]]))
    for G in Gs.tolist():
        cmax = min(1025, (1 << 21) // G + 1)  # at most 1024 bins, partitions below 2^21
        c = np.arange(0, cmax, dtype=np.int64)
        p = np.concatenate([((c * G)[:, None] + np.arange(-3, 4)[None, :]).ravel(),
                            ((c * G)[:, None] + rng.integers(0, G, (cmax, 8))).ravel()])
        p = p[(p >= 0) & (p < min(cmax * G, 1 << 21))]
        got_c, got_r = bin_f32(p, G)
        assert np.array_equal(got_c, p // G), G
        assert np.array_equal(got_r, p % G), G
