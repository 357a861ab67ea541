"""Exactness of the integer shortcuts the HIP kernels take, checked on the CPU.

The kernels replace two integer divisions with an f32 quotient estimate and
a 24-bit product (numpy float32 arithmetic rounds like the GPU's v_cvt_f32_u32
and v_mul_f32, to nearest even; v_cvt_u32_f32 truncates like astype):

- search_kernel.hip mod_f32: a mod d for a < 2^31, 512 <= d < 2^24, as the
  smallest of r, r + d, r - d with r = a - q*d (mod 2^32);
- partition_kernels.hip k_hash_route: the level-1 bin p / G and p mod G for
  G < 2^16, p < 2^21 partitions and p / G at most 1024 bins, with one
  correction step.

Both are compared with exact integer division, near every quotient boundary
and on random points.
"""
import numpy as np

U32 = np.uint32


def mod_f32(a: np.ndarray, d: int) -> np.ndarray:
    rd = np.float32(1.0 / d)  # (float)(1.0 / (double)d)
    q = (a.astype(np.float32) * rd).astype(np.uint64)
    r = (a.astype(np.uint64) - (q & 0xFFFFFF) * (d & 0xFFFFFF)) & 0xFFFFFFFF
    r1 = (r + d) & 0xFFFFFFFF
    r2 = (r - d) & 0xFFFFFFFF
    return np.minimum(np.minimum(r, r1), r2)


def bin_f32(p: np.ndarray, G: int):
    rG = np.float32(1.0) / np.float32(G)
    c = (p.astype(np.float32) * rG).astype(np.int64)
    r = p.astype(np.int64) - (c & 0xFFFFFF) * (G & 0xFFFFFF)
    lo, hi = r < 0, r >= G
    c = c - lo + hi
    r = r + lo * G - hi * G
    return c, r


def test_mod_f32_exact_near_boundaries_and_random():
    rng = np.random.default_rng(11)
    sizes = [512, 513, 1000, 2047, 2048, 2049, 2069, 2112, 3072, 3103, 4096, 4097, 6000, 8191, 8192, 65537,
             1 << 20, (1 << 24) - 1]
    for d in sizes:
        k = np.arange(0, (1 << 31) // d + 1, max(1, ((1 << 31) // d) // 200000), dtype=np.int64)
        base = k * d
        a = np.concatenate([base + o for o in (-2, -1, 0, 1, 2, d - 1)] +
                           [rng.integers(0, 1 << 31, 300000, dtype=np.int64)])
        a = a[(a >= 0) & (a < (1 << 31))]
        got = mod_f32(a, d)
        assert np.array_equal(got, (a % d).astype(np.uint64)), d


def test_mod_f32_exhaustive_small_range():
    # every a below 2^22 for a few table sizes around P = 2048
    a = np.arange(0, 1 << 22, dtype=np.int64)
    for d in (512, 2048, 2069, 3072):
        assert np.array_equal(mod_f32(a, d), (a % d).astype(np.uint64)), d


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
