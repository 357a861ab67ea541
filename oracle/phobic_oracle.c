/*
 * phobic_oracle.c -- CPU restatement of the PHOBIC construction path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Loaded by tests/, by the smoke entry point and by the benchmark's CPU
 * baseline leg.  Reference citations are relative to stefanfred/PHOBIC-GPU.
 *
 * Parity pins for each part live in tests/test_oracle.py:
 *   XXH3-128           python xxhash 3.8.1 (xxHash 0.8.2) vectors, committed
 *                      as tests/golden/xxh3_vectors.json
 *   fulcrums / B       the reference's own mphf_config.h + bucketers.h,
 *                      compiled by oracle/Makefile into oracle/_ref/
 *   hi32 multiply      exact 64-bit products
 *   mix32              invertibility (inverse built from the odd multiplier)
 *   bucket order       sort by (-size, -index) computed independently
 *   search             brute-force minimality + bijection on small inputs
 *   compact encoding   decode round trip + closed-form sizes
 */
#include "phobic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- XXH3 -- */
/* Restated from the xxHash 0.8.x XXH3 specification (Yann Collet, BSD-2).
 * The 192-byte default secret is a format constant of the algorithm. */

static const uint64_t XP32_1 = 0x9E3779B1u, XP32_2 = 0x85EBCA77u, XP32_3 = 0xC2B2AE3Du;
static const uint64_t XP64_1 = 0x9E3779B185EBCA87ull, XP64_2 = 0xC2B2AE3D27D4EB4Full,
                      XP64_3 = 0x165667B19E3779F9ull, XP64_4 = 0x85EBCA77C2B2AE63ull,
                      XP64_5 = 0x27D4EB2F165667C5ull;
static const uint64_t XMX1 = 0x165667919E3779F9ull, XMX2 = 0x9FB21C651E98DF25ull;

/* default secret, as 24 little-endian 64-bit words */
static const uint64_t kSecretWords[24] = {
    0xbe4ba423396cfeb8ull, 0x1cad21f72c81017cull, 0xdb979083e96dd4deull, 0x1f67b3b7a4a44072ull,
    0x78e5c0cc4ee679cbull, 0x2172ffcc7dd05a82ull, 0x8e2443f7744608b8ull, 0x4c263a81e69035e0ull,
    0xcb00c391bb52283cull, 0xa32e531b8b65d088ull, 0x4ef90da297486471ull, 0xd8acdea946ef1938ull,
    0x3f349ce33f76faa8ull, 0x1d4f0bc7c7bbdcf9ull, 0x3159b4cd4be0518aull, 0x647378d9c97e9fc8ull,
    0xc3ebd33483acc5eaull, 0xeb6313faffa081c5ull, 0x49daf0b751dd0d17ull, 0x9e68d429265516d3ull,
    0xfca1477d58be162bull, 0xce31d07ad1b8f88full, 0x280416958f3acb45ull, 0x7e404bbbcafbd7afull};

static uint8_t g_secret[192];

/* filled when the library is loaded, before any thread hashes */
__attribute__((constructor)) static void init_secret(void) {
  for (int i = 0; i < 24; i++)
    for (int b = 0; b < 8; b++) g_secret[8 * i + b] = (uint8_t)(kSecretWords[i] >> (8 * b));
}

static const uint8_t* default_secret(void) { return g_secret; }

static inline uint64_t rd64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}
static inline uint32_t rd32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline uint32_t bswap32(uint32_t x) {
  return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
}
static inline uint64_t bswap64(uint64_t x) {
  return ((uint64_t)bswap32((uint32_t)x) << 32) | bswap32((uint32_t)(x >> 32));
}
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

typedef struct {
  uint64_t lo, hi;
} u128p;

static inline u128p mul64x64(uint64_t a, uint64_t b) {
  unsigned __int128 r = (unsigned __int128)a * b;
  u128p o = {(uint64_t)r, (uint64_t)(r >> 64)};
  return o;
}
static inline uint64_t fold64(uint64_t a, uint64_t b) {
  u128p r = mul64x64(a, b);
  return r.lo ^ r.hi;
}
static inline uint64_t av_xxh3(uint64_t h) {
  h ^= h >> 37;
  h *= XMX1;
  return h ^ (h >> 32);
}
static inline uint64_t av_xxh64(uint64_t h) {
  h ^= h >> 33;
  h *= XP64_2;
  h ^= h >> 29;
  h *= XP64_3;
  return h ^ (h >> 32);
}

static inline uint64_t mix16(const uint8_t* in, const uint8_t* sec, uint64_t seed) {
  return fold64(rd64(in) ^ (rd64(sec) + seed), rd64(in + 8) ^ (rd64(sec + 8) - seed));
}

static inline u128p mix32(u128p acc, const uint8_t* a, const uint8_t* b, const uint8_t* sec,
                          uint64_t seed) {
  acc.lo += mix16(a, sec, seed);
  acc.lo ^= rd64(b) + rd64(b + 8);
  acc.hi += mix16(b, sec + 16, seed);
  acc.hi ^= rd64(a) + rd64(a + 8);
  return acc;
}

static void acc512(uint64_t* acc, const uint8_t* in, const uint8_t* sec) {
  for (int i = 0; i < 8; i++) {
    uint64_t v = rd64(in + 8 * i);
    uint64_t k = v ^ rd64(sec + 8 * i);
    acc[i ^ 1] += v;
    acc[i] += (uint64_t)(uint32_t)k * (k >> 32);
  }
}

static void scramble(uint64_t* acc, const uint8_t* sec) {
  for (int i = 0; i < 8; i++) {
    uint64_t a = acc[i];
    a ^= a >> 47;
    a ^= rd64(sec + 8 * i);
    acc[i] = a * XP32_1;
  }
}

static uint64_t merge_accs(const uint64_t* acc, const uint8_t* sec, uint64_t start) {
  uint64_t r = start;
  for (int i = 0; i < 4; i++) r += fold64(acc[2 * i] ^ rd64(sec + 16 * i), acc[2 * i + 1] ^ rd64(sec + 16 * i + 8));
  return av_xxh3(r);
}

void orc_xxh3_128(const void* data, size_t len, uint64_t seed, uint64_t* lo64, uint64_t* hi64) {
  const uint8_t* in = (const uint8_t*)data;
  const uint8_t* sec = default_secret();
  u128p h;
  if (len == 0) {
    h.lo = av_xxh64(seed ^ (rd64(sec + 64) ^ rd64(sec + 72)));
    h.hi = av_xxh64(seed ^ (rd64(sec + 80) ^ rd64(sec + 88)));
  } else if (len <= 3) {
    uint32_t c1 = in[0], c2 = in[len >> 1], c3 = in[len - 1];
    uint32_t cl = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
    uint32_t ch = rotl32(bswap32(cl), 13);
    uint64_t fl = (uint64_t)(rd32(sec) ^ rd32(sec + 4)) + seed;
    uint64_t fh = (uint64_t)(rd32(sec + 8) ^ rd32(sec + 12)) - seed;
    h.lo = av_xxh64((uint64_t)cl ^ fl);
    h.hi = av_xxh64((uint64_t)ch ^ fh);
  } else if (len <= 8) {
    uint64_t s = seed ^ ((uint64_t)bswap32((uint32_t)seed) << 32);
    uint64_t x = (uint64_t)rd32(in) + ((uint64_t)rd32(in + len - 4) << 32);
    uint64_t flip = (rd64(sec + 16) ^ rd64(sec + 24)) + s;
    h = mul64x64(x ^ flip, XP64_1 + ((uint64_t)len << 2));
    h.hi += h.lo << 1;
    h.lo ^= h.hi >> 3;
    h.lo ^= h.lo >> 35;
    h.lo *= XMX2;
    h.lo ^= h.lo >> 28;
    h.hi = av_xxh3(h.hi);
  } else if (len <= 16) {
    uint64_t fl = (rd64(sec + 32) ^ rd64(sec + 40)) - seed;
    uint64_t fh = (rd64(sec + 48) ^ rd64(sec + 56)) + seed;
    uint64_t a = rd64(in), b = rd64(in + len - 8);
    u128p m = mul64x64(a ^ b ^ fl, XP64_1);
    m.lo += (uint64_t)(len - 1) << 54;
    b ^= fh;
    m.hi += b + (uint64_t)(uint32_t)b * (XP32_2 - 1);
    m.lo ^= bswap64(m.hi);
    h = mul64x64(m.lo, XP64_2);
    h.hi += m.hi * XP64_2;
    h.lo = av_xxh3(h.lo);
    h.hi = av_xxh3(h.hi);
  } else if (len <= 128) {
    u128p acc = {(uint64_t)len * XP64_1, 0};
    if (len > 32) {
      if (len > 64) {
        if (len > 96) acc = mix32(acc, in + 48, in + len - 64, sec + 96, seed);
        acc = mix32(acc, in + 32, in + len - 48, sec + 64, seed);
      }
      acc = mix32(acc, in + 16, in + len - 32, sec + 32, seed);
    }
    acc = mix32(acc, in, in + len - 16, sec, seed);
    h.lo = av_xxh3(acc.lo + acc.hi);
    h.hi = 0 - av_xxh3(acc.lo * XP64_1 + acc.hi * XP64_4 + ((uint64_t)len - seed) * XP64_2);
  } else if (len <= 240) {
    u128p acc = {(uint64_t)len * XP64_1, 0};
    int rounds = (int)(len / 32);
    for (int i = 0; i < 4; i++) acc = mix32(acc, in + 32 * i, in + 32 * i + 16, sec + 32 * i, seed);
    acc.lo = av_xxh3(acc.lo);
    acc.hi = av_xxh3(acc.hi);
    for (int i = 4; i < rounds; i++)
      acc = mix32(acc, in + 32 * i, in + 32 * i + 16, sec + 3 + 32 * (i - 4), seed);
    acc = mix32(acc, in + len - 16, in + len - 32, sec + 136 - 17 - 16, 0 - seed);
    h.lo = av_xxh3(acc.lo + acc.hi);
    h.hi = 0 - av_xxh3(acc.lo * XP64_1 + acc.hi * XP64_4 + ((uint64_t)len - seed) * XP64_2);
  } else {
    uint8_t custom[192];
    const uint8_t* s = sec;
    if (seed != 0) {
      for (int i = 0; i < 12; i++) {
        uint64_t a = rd64(sec + 16 * i) + seed, b = rd64(sec + 16 * i + 8) - seed;
        for (int k = 0; k < 8; k++) {
          custom[16 * i + k] = (uint8_t)(a >> (8 * k));
          custom[16 * i + 8 + k] = (uint8_t)(b >> (8 * k));
        }
      }
      s = custom;
    }
    uint64_t acc[8] = {XP32_3, XP64_1, XP64_2, XP64_3, XP64_4, XP32_2, XP64_5, XP32_1};
    const size_t stripes_per_block = (192 - 64) / 8, block = 64 * stripes_per_block;
    size_t nb = (len - 1) / block;
    for (size_t b = 0; b < nb; b++) {
      for (size_t k = 0; k < stripes_per_block; k++) acc512(acc, in + b * block + 64 * k, s + 8 * k);
      scramble(acc, s + 192 - 64);
    }
    size_t last = ((len - 1) - block * nb) / 64;
    for (size_t k = 0; k < last; k++) acc512(acc, in + nb * block + 64 * k, s + 8 * k);
    acc512(acc, in + len - 64, s + 192 - 64 - 7);
    h.lo = merge_accs(acc, s + 11, (uint64_t)len * XP64_1);
    h.hi = merge_accs(acc, s + 192 - 64 - 11, ~((uint64_t)len * XP64_2));
  }
  *lo64 = h.lo;
  *hi64 = h.hi;
}

static inline orc_key key_of(uint64_t lo, uint64_t hi) {
  orc_key k = {(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)};
  return k;
}

/* hasher.hpp:42-48: XXH128(&val, 8, 0) on the key's native (little-endian) bytes. */
void orc_hash_u64(const uint64_t* keys, uint64_t n, uint64_t seed, orc_key* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint8_t b[8];
    for (int k = 0; k < 8; k++) b[k] = (uint8_t)(keys[i] >> (8 * k));
    uint64_t lo, hi;
    orc_xxh3_128(b, 8, seed, &lo, &hi);
    out[i] = key_of(lo, hi);
  }
}

/* hasher.hpp:33-39: XXH128(data(), size(), 0) of a string. */
void orc_hash_bytes(const uint8_t* bytes, const uint64_t* offsets, uint64_t n, uint64_t seed,
                    orc_key* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint64_t lo, hi;
    orc_xxh3_128(bytes + offsets[i], (size_t)(offsets[i + 1] - offsets[i]), seed, &lo, &hi);
    out[i] = key_of(lo, hi);
  }
}

/* -------------------------------------------------------- config ------ */

/* mphf_config.h:26-29: B = ceil(P / lambda). */
uint32_t orc_bucket_count(double lambda, uint32_t partition_size) {
  return (uint32_t)ceil((double)partition_size / lambda);
}

/* mphf_config.h:44-55 with OptBucketer (bucketers.h:28-39):
 * c = 0.2*lambda/sqrt(P); y(x) = (x + (1-x)ln(1-x))(1-c) + c x. */
void orc_fulcrums(double lambda, uint32_t partition_size, uint32_t* f) {
  uint32_t B = orc_bucket_count(lambda, partition_size);
  double c = 0.2 * lambda / sqrt((double)partition_size);
  double scale = (double)(uint32_t)(B << 16);
  f[0] = 0;
  for (int xi = 1; xi < ORC_FULCS - 1; xi++) {
    double x = (double)xi / (double)(ORC_FULCS - 1);
    double y = (x + (1.0 - x) * log(1.0 - x)) * (1.0 - c) + c * x;
    f[xi] = (uint32_t)(y * scale);
  }
  f[ORC_FULCS - 1] = B << 16;
}

void orc_sizes(uint64_t n, double lambda, uint32_t partition_size, uint32_t* np, uint32_t* B) {
  *np = (uint32_t)((n + partition_size - 1) / partition_size);
  *B = orc_bucket_count(lambda, partition_size);
}

/* ------------------------------------------------------ key mapping --- */

/* key_mapping.glsl:9-29: high 32 bits of a 32x32 product. */
uint32_t orc_hi32_mul(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

/* key_mapping.glsl:43-45 */
uint32_t orc_partition_of(uint32_t partitioner, uint32_t np) { return orc_hi32_mul(partitioner, np); }

/* key_mapping.glsl:34-41 (host twin mphf.hpp:42-49): note fulcs[i] is
 * weighted by the fractional part and fulcs[i+1] by its complement. */
uint32_t orc_bucket_of(uint32_t bucketer, const uint32_t* fulcs) {
  uint32_t part = bucketer * (uint32_t)(ORC_FULCS - 1);
  uint32_t idx = orc_hi32_mul(bucketer, ORC_FULCS - 1);
  uint32_t v1 = orc_hi32_mul(fulcs[idx], part);
  uint32_t v2 = orc_hi32_mul(fulcs[idx + 1], 0xFFFFFFFFu - part);
  return (v1 + v2) >> 16;
}

/* hashing.glsl:11-17 */
uint32_t orc_mix32(uint32_t key, uint32_t pilot) {
  key ^= pilot;
  key = ((key >> 16) ^ key) * 0x45d9f3bu;
  key = ((key >> 16) ^ key) * 0x45d9f3bu;
  return (key >> 16) ^ key;
}

/* search.comp:31-33,47,62: initial position of a key for a pilot. */
static inline uint32_t init_pos(const orc_key* k, uint32_t pilot, uint32_t ps) {
  return (orc_mix32(k->lower1, orc_mix32(k->lower2, pilot)) >> 1) % ps;
}

/* ------------------------------------------------------------ search -- */

/* Search one partition.  keys[0..ps) grouped by bucket in processing order
 * (bucket_list[j] = relative bucket, bucket_start[j]/bucket_size[j] = its
 * slice).  Implements search.comp:105-175: for every bucket in order, the
 * smallest V = d + ps*pilot whose positions (init_pos + d) mod ps are
 * pairwise distinct and untaken; then mark them taken. */
static int search_partition(const orc_key* keys, uint32_t ps, uint32_t nbk, const uint32_t* bucket_list,
                            const uint32_t* bucket_start, const uint32_t* bucket_size, uint32_t* pilot_of,
                            uint64_t** taken_out) {
  /* ps here is the table size m of the partition (= key count when alpha = 1) */
  uint64_t words = ((uint64_t)ps + 63) / 64;
  uint64_t* taken = (uint64_t*)calloc(words ? words : 1, 8);
  uint64_t* local = (uint64_t*)calloc(words ? words : 1, 8);
  uint32_t pos[ORC_MAX_BUCKET];
  int status = ORC_OK;
  for (uint32_t j = 0; j < nbk && status == ORC_OK; j++) {
    const orc_key* kb = keys + bucket_start[j];
    uint32_t s = bucket_size[j];
    int found = 0;
    for (uint64_t pilot = 0; !found; pilot++) {
      if ((uint64_t)ps * pilot > 0x7FFFFFFFull) {
        status = ORC_E_PILOT_OVERFLOW;
        break;
      }
      int coll = 0;
      for (uint32_t k = 0; k < s; k++) {
        pos[k] = init_pos(kb + k, (uint32_t)pilot, ps);
        uint64_t m = 1ull << (pos[k] & 63);
        if (local[pos[k] >> 6] & m) coll = 1;
        local[pos[k] >> 6] |= m;
      }
      for (uint32_t k = 0; k < s; k++) local[pos[k] >> 6] = 0;
      if (coll) {
        if (pilot == 0) {
          for (uint32_t a = 0; a < s && !status; a++)
            for (uint32_t b = a + 1; b < s; b++)
              if (kb[a].lower1 == kb[b].lower1 && kb[a].lower2 == kb[b].lower2) {
                status = ORC_E_DUPLICATE_KEYS;
                break;
              }
          if (status) break;
        }
        continue;
      }
      for (uint32_t d = 0; d < ps; d++) {
        if ((uint64_t)d + (uint64_t)ps * pilot > 0x7FFFFFFFull) break;
        uint32_t k = 0;
        for (; k < s; k++) {
          uint32_t q = pos[k] + d;
          if (q >= ps) q -= ps;
          if (taken[q >> 6] & (1ull << (q & 63))) break;
        }
        if (k == s) {
          for (k = 0; k < s; k++) {
            uint32_t q = pos[k] + d;
            if (q >= ps) q -= ps;
            taken[q >> 6] |= 1ull << (q & 63);
          }
          pilot_of[j] = d + ps * (uint32_t)pilot;
          found = 1;
          break;
        }
      }
    }
  }
  if (taken_out)
    *taken_out = taken;
  else
    free(taken);
  free(local);
  return status;
}

/* Table size of a partition with ps keys at load factor alpha = a_ppm / 1e6
 * (this library's extension; the reference only has alpha = 1):
 * m = ceil(ps * 1e6 / a_ppm). */
uint32_t orc_table_size(uint32_t ps, uint32_t a_ppm) {
  if (a_ppm >= 1000000u || a_ppm == 0) return ps;
  return (uint32_t)(((uint64_t)ps * 1000000u + a_ppm - 1) / a_ppm);
}

uint64_t orc_remap_capacity(uint64_t n, uint32_t np, uint32_t a_ppm) {
  if (a_ppm >= 1000000u || a_ppm == 0) return 0;
  return (n * (1000000u - a_ppm) + a_ppm - 1) / a_ppm + np;
}

/* Several errors in one build report the smallest status code. */
static int merge_err(int a, int b) {
  if (!a) return b;
  if (!b) return a;
  return a < b ? a : b;
}

int orc_build_keys(const orc_key* keys, uint64_t n, double lambda, uint32_t P, uint32_t* pilots,
                   uint32_t* offsets, uint32_t* bucket_sizes_out, uint32_t* order_out) {
  return orc_build_keys_ex(keys, n, lambda, P, 1000000u, pilots, offsets, NULL, NULL, bucket_sizes_out, order_out);
}

/* Build with load factor alpha = a_ppm / 1e6.  For alpha < 1 partition p
 * searches a table of m_p = orc_table_size(ps_p) slots; the used slots q in
 * [ps_p, m_p) are remapped, in increasing order, onto the free slots of
 * [0, ps_p), also in increasing order: remap[remap_offsets[p] + q - ps_p] is
 * the slot q stands for (0 for unused q).  remap_offsets[np] = entries. */
static int build_keys_core(const orc_key* keys, uint64_t n, uint32_t P, uint32_t B, const uint32_t* fulcs_in,
                           uint32_t a_ppm, uint32_t* pilots, uint32_t* offsets, uint32_t* remap,
                           uint32_t* remap_offsets, uint32_t* bucket_sizes_out, uint32_t* order_out);

int orc_build_keys_ex(const orc_key* keys, uint64_t n, double lambda, uint32_t P, uint32_t a_ppm, uint32_t* pilots,
                      uint32_t* offsets, uint32_t* remap, uint32_t* remap_offsets, uint32_t* bucket_sizes_out,
                      uint32_t* order_out) {
  if (P == 0 || !(lambda > 0)) return ORC_E_INVALID;
  uint32_t fulcs[ORC_FULCS];
  orc_fulcrums(lambda, P, fulcs);
  return build_keys_core(keys, n, P, orc_bucket_count(lambda, P), fulcs, a_ppm, pilots, offsets, remap,
                         remap_offsets, bucket_sizes_out, order_out);
}

/* A custom Bucketer's B and fulcrum table (mphf_config.h:34-42). */
int orc_build_keys_table(const orc_key* keys, uint64_t n, uint32_t P, uint32_t B, const uint32_t* fulcs,
                         uint32_t a_ppm, uint32_t* pilots, uint32_t* offsets, uint32_t* remap,
                         uint32_t* remap_offsets) {
  if (!fulcs) return ORC_E_INVALID;
  for (int i = 0; i < ORC_FULCS; i++)
    if ((uint64_t)fulcs[i] > ((uint64_t)B << 16)) return ORC_E_INVALID;
  return build_keys_core(keys, n, P, B, fulcs, a_ppm, pilots, offsets, remap, remap_offsets, NULL, NULL);
}

static int build_keys_core(const orc_key* keys, uint64_t n, uint32_t P, uint32_t B, const uint32_t* fulcs_in,
                           uint32_t a_ppm, uint32_t* pilots, uint32_t* offsets, uint32_t* remap,
                           uint32_t* remap_offsets, uint32_t* bucket_sizes_out, uint32_t* order_out) {
  if (P == 0 || n > 0xFFFFFFFFull || a_ppm == 0 || a_ppm > 1000000u) return ORC_E_INVALID;
  const int use_remap = a_ppm < 1000000u;
  if (use_remap && (!remap || !remap_offsets)) return ORC_E_INVALID;
  const uint32_t np = (uint32_t)((n + P - 1) / P);
  if (B == 0 || B > 65535) return ORC_E_INVALID;
  uint32_t* fulcs = (uint32_t*)malloc(ORC_FULCS * 4);
  memcpy(fulcs, fulcs_in, ORC_FULCS * 4);
  uint64_t NB = (uint64_t)np * B;
  uint32_t* gb = (uint32_t*)malloc((n ? n : 1) * 4);
  uint32_t* cnt = (uint32_t*)calloc(NB + 1, 4);
  /* bucket_sizes.comp:21-35: bucket of every key and per-bucket counts */
  for (uint64_t i = 0; i < n; i++) {
    uint32_t p = orc_partition_of(keys[i].partitioner, np);
    uint32_t b = orc_bucket_of(keys[i].bucketer, fulcs);
    gb[i] = p * B + b;
    cnt[gb[i]]++;
  }
  if (bucket_sizes_out) memcpy(bucket_sizes_out, cnt, NB * 4);
  int status = ORC_OK;
  uint32_t max_ps = P + P / 2; /* mphf_config.h:61-64 partitionMaxSize() */
  offsets[0] = 0;
  for (uint32_t p = 0; p < np; p++) {
    uint64_t ps = 0;
    for (uint32_t b = 0; b < B; b++) {
      if (cnt[(uint64_t)p * B + b] > ORC_MAX_BUCKET) status = merge_err(status, ORC_E_BUCKET_OVERFLOW);
      ps += cnt[(uint64_t)p * B + b];
    }
    if (ps > max_ps) status = merge_err(status, ORC_E_PARTITION_OVERFLOW);
    offsets[p + 1] = offsets[p] + (uint32_t)ps;
  }
  if (status) {
    free(fulcs);
    free(gb);
    free(cnt);
    return status;
  }
  /* group keys by partition and bucket, in processing order:
   * sort_sum.comp:156-172 bins buckets by size (largest first); within a
   * size, the serial loop runs index B-1 .. 0, so larger index first. */
  uint32_t* order = (uint32_t*)malloc(NB * 4 + 4);
  uint32_t* bstart = (uint32_t*)malloc(NB * 4 + 4); /* global start of each bucket */
  uint32_t* nbk = (uint32_t*)calloc(np + 1, 4);
  for (uint32_t p = 0; p < np; p++) {
    uint32_t* ord = order + (uint64_t)p * B;
    const uint32_t* c = cnt + (uint64_t)p * B;
    uint32_t m = 0;
    for (uint32_t size = ORC_MAX_BUCKET; size >= 1; size--)
      for (int64_t b = (int64_t)B - 1; b >= 0; b--)
        if (c[b] == size) ord[m++] = (uint32_t)b;
    nbk[p] = m;
    uint32_t at = offsets[p];
    for (uint32_t j = 0; j < m; j++) {
      bstart[(uint64_t)p * B + ord[j]] = at;
      at += c[ord[j]];
    }
    for (uint32_t j = m; j < B; j++) ord[j] = 0xFFFFFFFFu;
  }
  if (order_out) memcpy(order_out, order, NB * 4);
  /* redistribute_keys.comp:21-53 (order inside a bucket is irrelevant) */
  orc_key* sorted = (orc_key*)malloc((n ? n : 1) * sizeof(orc_key));
  uint32_t* fill = (uint32_t*)calloc(NB + 1, 4);
  for (uint64_t i = 0; i < n; i++) sorted[bstart[gb[i]] + fill[gb[i]]++] = keys[i];
  free(fill);
  free(gb);
  memset(pilots, 0, NB * 4); /* empty buckets: pilot 0 */
  if (use_remap) {
    remap_offsets[0] = 0;
    for (uint32_t p = 0; p < np; p++) {
      const uint32_t ps = offsets[p + 1] - offsets[p];
      remap_offsets[p + 1] = remap_offsets[p] + (orc_table_size(ps, a_ppm) - ps);
    }
  }
  int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < (int64_t)np; p++) {
    uint32_t m = nbk[p];
    const uint32_t* ord = order + (uint64_t)p * B;
    uint32_t* st = (uint32_t*)malloc((m ? m : 1) * 4);
    uint32_t* sz = (uint32_t*)malloc((m ? m : 1) * 4);
    uint32_t* pv = (uint32_t*)malloc((m ? m : 1) * 4);
    uint32_t at = 0;
    for (uint32_t j = 0; j < m; j++) {
      st[j] = at;
      sz[j] = cnt[(uint64_t)p * B + ord[j]];
      at += sz[j];
    }
    const uint32_t ps = offsets[p + 1] - offsets[p];
    const uint32_t mt = orc_table_size(ps, a_ppm);
    uint64_t* taken = NULL;
    int s = search_partition(sorted + offsets[p], mt, m, ord, st, sz, pv, use_remap ? &taken : NULL);
    if (s) {
#pragma omp critical
      err = merge_err(err, s);
    }
    if (use_remap && !s) {
      /* k-th used slot of [ps, mt) -> k-th free slot of [0, ps) */
      uint32_t* out = remap + remap_offsets[p];
      uint32_t hole = 0;
      for (uint32_t q = ps; q < mt; q++) {
        out[q - ps] = 0;
        if (taken[q >> 6] & (1ull << (q & 63))) {
          while (taken[hole >> 6] & (1ull << (hole & 63))) hole++;
          out[q - ps] = hole++;
        }
      }
    }
    free(taken);
    for (uint32_t j = 0; j < m; j++) pilots[(uint64_t)p + (uint64_t)ord[j] * np] = pv[j];
    free(st);
    free(sz);
    free(pv);
  }
  free(sorted);
  free(order);
  free(bstart);
  free(nbk);
  free(cnt);
  free(fulcs);
  return err;
}

/* ------------------------------------------- partitions of a key stream -- */

/* splitmix64 output i+1 from state `seed` (Steele, Lea, Flood 2014). */
static inline uint64_t splitmix_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline orc_key hash_one_u64(uint64_t v, uint64_t seed) {
  uint8_t b[8];
  for (int k = 0; k < 8; k++) b[k] = (uint8_t)(v >> (8 * k));
  uint64_t lo, hi;
  orc_xxh3_128(b, 8, seed, &lo, &hi);
  return key_of(lo, hi);
}

/* One partition built on its own from its ps keys (any order): the same
 * steps orc_build_keys_ex takes per partition -- bucket counts
 * (bucket_sizes.comp:21-35), processing order size descending then index
 * descending (sort_sum.comp:156-172), keys grouped by bucket
 * (redistribute_keys.comp:21-53), the search (search.comp:105-175) over
 * orc_table_size(ps) slots and, for a_ppm < 1e6, the remap of the used slots
 * of [ps, m) onto the free slots of [0, ps).  pil[b] receives bucket b's
 * pilot (0 when empty); remap receives m - ps entries. */
static int build_partition(const orc_key* keys, uint32_t ps, uint32_t B, const uint32_t* fulcs, uint32_t a_ppm,
                           uint32_t* pil, uint32_t* remap) {
  uint32_t* cnt = (uint32_t*)calloc(B, 4);
  uint32_t* kb = (uint32_t*)malloc((ps ? ps : 1) * 4);
  int status = ORC_OK;
  for (uint32_t i = 0; i < ps; i++) {
    kb[i] = orc_bucket_of(keys[i].bucketer, fulcs);
    cnt[kb[i]]++;
  }
  for (uint32_t b = 0; b < B; b++)
    if (cnt[b] > ORC_MAX_BUCKET) status = ORC_E_BUCKET_OVERFLOW;
  memset(pil, 0, (size_t)B * 4);
  if (status) {
    free(cnt);
    free(kb);
    return status;
  }
  uint32_t* ord = (uint32_t*)malloc(B * 4);
  uint32_t* st = (uint32_t*)malloc(B * 4);
  uint32_t* sz = (uint32_t*)malloc(B * 4);
  uint32_t* pv = (uint32_t*)malloc(B * 4);
  uint32_t* start_of = (uint32_t*)malloc(B * 4);
  uint32_t m = 0, at = 0;
  for (uint32_t size = ORC_MAX_BUCKET; size >= 1; size--)
    for (int64_t b = (int64_t)B - 1; b >= 0; b--)
      if (cnt[b] == size) ord[m++] = (uint32_t)b;
  for (uint32_t j = 0; j < m; j++) {
    st[j] = at;
    sz[j] = cnt[ord[j]];
    start_of[ord[j]] = at;
    at += sz[j];
  }
  orc_key* grouped = (orc_key*)malloc((ps ? ps : 1) * sizeof(orc_key));
  for (uint32_t i = 0; i < ps; i++) grouped[start_of[kb[i]]++] = keys[i];
  const uint32_t mt = orc_table_size(ps, a_ppm);
  const int use_remap = a_ppm < 1000000u;
  uint64_t* taken = NULL;
  status = search_partition(grouped, mt, m, ord, st, sz, pv, use_remap ? &taken : NULL);
  if (!status) {
    for (uint32_t j = 0; j < m; j++) pil[ord[j]] = pv[j];
    if (use_remap) {
      uint32_t hole = 0;
      for (uint32_t q = ps; q < mt; q++) {
        remap[q - ps] = 0;
        if (taken[q >> 6] & (1ull << (q & 63))) {
          while (taken[hole >> 6] & (1ull << (hole & 63))) hole++;
          remap[q - ps] = hole++;
        }
      }
    }
  }
  free(taken);
  free(grouped);
  free(start_of);
  free(pv);
  free(sz);
  free(st);
  free(ord);
  free(kb);
  free(cnt);
  return status;
}

/* Per-partition key counts and largest bucket of the key stream
 * splitmix64(kseed) outputs 1..n hashed with XXH3 seed hseed
 * (hasher.hpp:42-48), without holding the keys: one pass with atomic
 * counters.  max_bucket may be NULL. */
int orc_scan_splitmix(uint64_t kseed, uint64_t n, uint64_t hseed, double lambda, uint32_t P, uint32_t* counts,
                      uint16_t* max_bucket) {
  if (P == 0 || !(lambda > 0) || n > 0xFFFFFFFFull || !counts) return ORC_E_INVALID;
  uint32_t np, B;
  orc_sizes(n, lambda, P, &np, &B);
  uint32_t fulcs[ORC_FULCS];
  orc_fulcrums(lambda, P, fulcs);
  memset(counts, 0, (size_t)np * 4);
  uint16_t* bc = max_bucket ? (uint16_t*)calloc((size_t)np * B + 1, 2) : NULL;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const orc_key k = hash_one_u64(splitmix_at(kseed, (uint64_t)i), hseed);
    const uint32_t p = orc_partition_of(k.partitioner, np);
    __atomic_fetch_add(&counts[p], 1u, __ATOMIC_RELAXED);
    if (bc) {
      const uint64_t g = (uint64_t)p * B + orc_bucket_of(k.bucketer, fulcs);
      uint16_t c = __atomic_load_n(&bc[g], __ATOMIC_RELAXED);
      while (c < 0xFFFFu && !__atomic_compare_exchange_n(&bc[g], &c, (uint16_t)(c + 1), 1, __ATOMIC_RELAXED,
                                                        __ATOMIC_RELAXED)) {
      }
    }
  }
  if (bc) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)np; p++) {
      uint16_t mx = 0;
      for (uint32_t b = 0; b < B; b++) mx = bc[(uint64_t)p * B + b] > mx ? bc[(uint64_t)p * B + b] : mx;
      max_bucket[p] = mx;
    }
    free(bc);
  }
  return ORC_OK;
}

/* Builds only the partitions parts[0..nparts) (distinct, any order) of the
 * key stream of orc_scan_splitmix: pilots_sel[j*B + b] is the pilot of
 * bucket b of partition parts[j]; counts_sel[j] its key count; for
 * a_ppm < 1e6 remap_sel receives, for j in list order, the
 * orc_table_size(counts_sel[j]) - counts_sel[j] remap entries of partition
 * parts[j] (remap_sel_cap entries at most).  Each partition depends only on
 * its own keys, so these equal the corresponding rows of a full build. */
int orc_build_splitmix_parts(uint64_t kseed, uint64_t n, uint64_t hseed, double lambda, uint32_t P, uint32_t a_ppm,
                             const uint32_t* parts, uint32_t nparts, uint32_t* pilots_sel, uint32_t* counts_sel,
                             uint32_t* remap_sel, uint64_t remap_sel_cap) {
  if (P == 0 || !(lambda > 0) || n > 0xFFFFFFFFull || a_ppm == 0 || a_ppm > 1000000u) return ORC_E_INVALID;
  uint32_t np, B;
  orc_sizes(n, lambda, P, &np, &B);
  if (B == 0 || B > 65535) return ORC_E_INVALID;
  const int use_remap = a_ppm < 1000000u;
  if (use_remap && !remap_sel) return ORC_E_INVALID;
  int32_t* slot = (int32_t*)malloc((size_t)(np ? np : 1) * 4);
  for (uint32_t p = 0; p < np; p++) slot[p] = -1;
  for (uint32_t j = 0; j < nparts; j++) {
    if (parts[j] >= np || slot[parts[j]] >= 0) {
      free(slot);
      return ORC_E_INVALID;
    }
    slot[parts[j]] = (int32_t)j;
  }
  uint32_t fulcs[ORC_FULCS];
  orc_fulcrums(lambda, P, fulcs);
  /* pass 1: keys per listed partition; pass 2: the keys themselves */
  uint32_t* cnt = (uint32_t*)calloc((size_t)nparts + 1, 4);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const orc_key k = hash_one_u64(splitmix_at(kseed, (uint64_t)i), hseed);
    const int32_t j = slot[orc_partition_of(k.partitioner, np)];
    if (j >= 0) __atomic_fetch_add(&cnt[j], 1u, __ATOMIC_RELAXED);
  }
  uint64_t* start = (uint64_t*)malloc(((size_t)nparts + 1) * 8);
  uint64_t total = 0, rtotal = 0;
  int status = ORC_OK;
  for (uint32_t j = 0; j < nparts; j++) {
    start[j] = total;
    total += cnt[j];
    counts_sel[j] = cnt[j];
    if (cnt[j] > P + P / 2) status = merge_err(status, ORC_E_PARTITION_OVERFLOW); /* mphf_config.h:61-64 */
    rtotal += orc_table_size(cnt[j], a_ppm) - cnt[j];
  }
  start[nparts] = total;
  if (use_remap && rtotal > remap_sel_cap) status = ORC_E_INVALID;
  if (status) {
    free(start);
    free(cnt);
    free(slot);
    return status;
  }
  orc_key* keys = (orc_key*)malloc((size_t)(total ? total : 1) * sizeof(orc_key));
  uint32_t* fill = (uint32_t*)calloc((size_t)nparts + 1, 4);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const orc_key k = hash_one_u64(splitmix_at(kseed, (uint64_t)i), hseed);
    const int32_t j = slot[orc_partition_of(k.partitioner, np)];
    if (j >= 0) keys[start[j] + __atomic_fetch_add(&fill[j], 1u, __ATOMIC_RELAXED)] = k;
  }
  uint64_t* rstart = (uint64_t*)malloc(((size_t)nparts + 1) * 8);
  rtotal = 0;
  for (uint32_t j = 0; j < nparts; j++) {
    rstart[j] = rtotal;
    rtotal += orc_table_size(cnt[j], a_ppm) - cnt[j];
  }
  int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t j = 0; j < (int64_t)nparts; j++) {
    const int s = build_partition(keys + start[j], cnt[j], B, fulcs, a_ppm, pilots_sel + (uint64_t)j * B,
                                  use_remap ? remap_sel + rstart[j] : NULL);
    if (s) {
#pragma omp critical
      err = merge_err(err, s);
    }
  }
  free(rstart);
  free(fill);
  free(keys);
  free(start);
  free(cnt);
  free(slot);
  return err;
}

/* mphf.hpp:32-40,77-86 */
void orc_query_keys(const orc_key* keys, uint64_t n, const uint32_t* pilots, const uint32_t* offsets,
                    uint32_t np, uint32_t B, const uint32_t* fulcs, uint32_t* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t p = orc_partition_of(keys[i].partitioner, np);
    uint32_t b = orc_bucket_of(keys[i].bucketer, fulcs);
    uint32_t V = pilots[(uint64_t)p + (uint64_t)b * np];
    uint32_t off = offsets[p], ps = offsets[p + 1] - offsets[p];
    uint32_t hp = V / ps;
    uint32_t hv = orc_mix32(keys[i].lower1, orc_mix32(keys[i].lower2, hp)) >> 1;
    out[i] = off + (hv + V) % ps;
  }
}

/* Query for an alpha < 1 build (remap may be NULL when a_ppm = 1e6). */
void orc_query_keys_ex(const orc_key* keys, uint64_t n, const uint32_t* pilots, const uint32_t* offsets,
                       uint32_t np, uint32_t B, const uint32_t* fulcs, uint32_t a_ppm, const uint32_t* remap,
                       const uint32_t* remap_offsets, uint32_t* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t p = orc_partition_of(keys[i].partitioner, np);
    uint32_t b = orc_bucket_of(keys[i].bucketer, fulcs);
    uint32_t V = pilots[(uint64_t)p + (uint64_t)b * np];
    uint32_t off = offsets[p], ps = offsets[p + 1] - offsets[p];
    uint32_t m = orc_table_size(ps, a_ppm);
    uint32_t hv = orc_mix32(keys[i].lower1, orc_mix32(keys[i].lower2, V / m)) >> 1;
    uint32_t pos = (hv + V) % m;
    if (pos >= ps) pos = remap[remap_offsets[p] + pos - ps];
    out[i] = off + pos;
  }
}

/* ----------------------------------------------------------- encoders -- */

uint64_t orc_compact_words(uint64_t n, uint64_t width) { return (n * width + 63) / 64 + 1; }

/* compact_vector.hpp:325-330: width = max==0 ? 1 : ceil(log2(max+1)) */
uint64_t orc_compact_width(const uint64_t* values, uint64_t n) {
  uint64_t mx = 0;
  for (uint64_t i = 0; i < n; i++)
    if (values[i] > mx) mx = values[i];
  uint64_t w = 0;
  while (w < 64 && (mx >> w) != 0) w++;
  return w ? w : 1;
}

/* compact_vector.hpp:243-265: LSB-first packing into 64-bit words. */
uint64_t orc_compact_encode(const uint64_t* values, uint64_t n, uint64_t* words, uint64_t* width_out) {
  uint64_t w = orc_compact_width(values, n);
  uint64_t nw = orc_compact_words(n, w);
  memset(words, 0, nw * 8);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t pos = i * w, blk = pos >> 6, sh = pos & 63;
    words[blk] |= values[i] << sh;
    if (sh + w > 64) words[blk + 1] |= values[i] >> (64 - sh);
  }
  *width_out = w;
  return nw;
}

uint64_t orc_compact_access(const uint64_t* words, uint64_t w, uint64_t i) {
  uint64_t pos = i * w, blk = pos >> 6, sh = pos & 63;
  uint64_t mask = w == 64 ? ~0ull : ((1ull << w) - 1);
  uint64_t v = words[blk] >> sh;
  if (sh + w > 64) v |= words[blk + 1] << (64 - sh);
  return v & mask;
}

static uint8_t* put64(uint8_t* o, uint64_t v) {
  if (o)
    for (int i = 0; i < 8; i++) o[i] = (uint8_t)(v >> (8 * i));
  return o ? o + 8 : NULL;
}

/* compact_vector::visit order (compact_vector.hpp:276-282): size, width,
 * mask, word count, words. */
static uint64_t put_compact(uint8_t** o, const uint64_t* vals, uint64_t n) {
  uint64_t w = orc_compact_width(vals, n);
  uint64_t nw = orc_compact_words(n, w);
  if (*o) {
    uint64_t* words = (uint64_t*)calloc(nw, 8);
    orc_compact_encode(vals, n, words, &w);
    *o = put64(*o, n);
    *o = put64(*o, w);
    *o = put64(*o, w == 64 ? ~0ull : ((1ull << w) - 1));
    *o = put64(*o, nw);
    for (uint64_t i = 0; i < nw; i++) *o = put64(*o, words[i]);
    free(words);
  }
  return 32 + 8 * nw;
}

/* Rice coding of a column (rice_sequence.hpp:14-46): parameter l from the
 * mean (Kiely's rule, double precision as in optimalL, 0 for an all-zero
 * column), low l bits in a compact vector of width l, v >> l in unary (zeros
 * then a one). */
uint64_t orc_rice_param(const uint64_t* v, uint64_t n) {
  uint64_t sum = 0;
  for (uint64_t i = 0; i < n; i++) sum += v[i];
  if (sum == 0) return 0;
  const double p = (double)n / ((double)sum + (double)n);
  const double gold = log2(2.0 / (sqrt(5.0) + 1.0));
  const double l = ceil(log2(gold / log2(1.0 - p)));
  return l > 0 ? (uint64_t)l : 0;
}

/* A serialised Rice column: the unary bit vector (size, word count, words),
 * its select index (darray.hpp:13-118: one count, blocks of 4096 ones as the
 * first one's position or, when the block's ones span 2^16 bits or more,
 * -1 - the index of its explicit positions; 16-bit offsets of every 64th one
 * from the block start (0xFFFF in explicit blocks), packed four per word;
 * the explicit positions) and the low bits (compact vector of width l).
 * Returns the serialised byte count; *ref_bytes receives the reference's
 * byte accounting (rice_sequence::bytes(), where 16-bit entries count 2). */
static uint64_t put_rice(uint8_t** o, const uint64_t* v, uint64_t n, uint64_t* ref_bytes) {
  const uint64_t l = orc_rice_param(v, n);
  uint64_t* pos = (uint64_t*)malloc((n ? n : 1) * 8); /* position of value i's one */
  uint64_t bits = 0;
  for (uint64_t i = 0; i < n; i++) {
    bits += (l >= 64 ? 0 : v[i] >> l) + 1;
    pos[i] = bits - 1;
  }
  const uint64_t hw = (bits + 63) / 64;
  const uint64_t nb = (n + 4095) / 4096;
  uint64_t ns = 0, nsp = 0;
  for (uint64_t b = 0; b < nb; b++) {
    const uint64_t lo = 4096 * b, hi = lo + 4096 < n ? lo + 4096 : n;
    ns += (hi - lo + 63) / 64;
    if (pos[hi - 1] - pos[lo] >= 65536) nsp += hi - lo;
  }
  const uint64_t lw = orc_compact_words(n, l);
  if (*o) {
    uint64_t* words = (uint64_t*)calloc(hw ? hw : 1, 8);
    for (uint64_t i = 0; i < n; i++) words[pos[i] >> 6] |= 1ull << (pos[i] & 63);
    *o = put64(*o, bits);
    *o = put64(*o, hw);
    for (uint64_t i = 0; i < hw; i++) *o = put64(*o, words[i]);
    free(words);
    *o = put64(*o, n);
    *o = put64(*o, nb);
    uint16_t* subs = (uint16_t*)calloc(ns ? ns : 1, 2);
    uint64_t* sparse = (uint64_t*)malloc((nsp ? nsp : 1) * 8);
    uint64_t si = 0, spi = 0;
    for (uint64_t b = 0; b < nb; b++) {
      const uint64_t lo = 4096 * b, hi = lo + 4096 < n ? lo + 4096 : n;
      const int dense = pos[hi - 1] - pos[lo] < 65536;
      *o = put64(*o, dense ? pos[lo] : (uint64_t)(-(int64_t)spi - 1));
      for (uint64_t i = lo; i < hi; i += 64) subs[si++] = dense ? (uint16_t)(pos[i] - pos[lo]) : 0xFFFF;
      if (!dense)
        for (uint64_t i = lo; i < hi; i++) sparse[spi++] = pos[i];
    }
    *o = put64(*o, ns);
    for (uint64_t i = 0; i < ns; i += 4) {
      uint64_t w = 0;
      for (uint64_t k = 0; k < 4 && i + k < ns; k++) w |= (uint64_t)subs[i + k] << (16 * k);
      *o = put64(*o, w);
    }
    *o = put64(*o, nsp);
    for (uint64_t i = 0; i < nsp; i++) *o = put64(*o, sparse[i]);
    free(subs);
    free(sparse);
    /* low bits: size, width, mask, word count, words */
    uint64_t* low = (uint64_t*)calloc(lw, 8);
    const uint64_t mask = l >= 64 ? ~0ull : ((1ull << l) - 1);
    for (uint64_t i = 0; i < n && l; i++) {
      const uint64_t x = v[i] & mask, p2 = i * l, blk = p2 >> 6, sh = p2 & 63;
      low[blk] |= x << sh;
      if (sh + l > 64) low[blk + 1] |= x >> (64 - sh);
    }
    *o = put64(*o, n);
    *o = put64(*o, l);
    *o = put64(*o, mask);
    *o = put64(*o, lw);
    for (uint64_t i = 0; i < lw; i++) *o = put64(*o, low[i]);
    free(low);
  }
  free(pos);
  *ref_bytes = (16 + 8 * hw) + (8 + (8 + 8 * nb) + (8 + 2 * ns) + (8 + 8 * nsp)) + (32 + 8 * lw);
  return (16 + 8 * hw) + (8 + (8 + 8 * nb) + (8 + 8 * ((ns + 3) / 4)) + (8 + 8 * nsp)) + (32 + 8 * lw);
}

/* One base-encoded sequence (base 0 compact, 1 rice); returns the serialised
 * bytes and adds the bytes num_bits() counts to *bpk_bytes. */
static uint64_t put_seq(uint8_t** o, const uint64_t* v, uint64_t n, uint32_t base, uint64_t* bpk_bytes) {
  if (base == 1) {
    uint64_t rb = 0;
    const uint64_t t = put_rice(o, v, n, &rb);
    *bpk_bytes += rb;
    return t;
  }
  const uint64_t cb = put_compact(o, v, n);
  *bpk_bytes += cb;
  return cb;
}

/* Serialised MPHF ("PHOBMPH2"): header (n, P, np, B, seed, encoding, load
 * factor ppm), fulcrums, the pilots, then the partition offsets.  Pilots:
 * B columns (interleaved_encoder.hpp:9-15: column j = pilots[j*np ..
 * j*np+np)) as compact vectors (encoding 0, FastQueryMphf) or Rice sequences
 * (encoding 1, SmallSpaceMphf); np then one sequence of all np*B pilots
 * (encodings 2, 3); or the dual split point then columns of two bases
 * (encodings 4..7).  Offsets: P then linear_diff_encoder
 * (linear_diff_encoder.hpp:14-28: v_i = offsets[i] - i*P, stored as
 * (|v|<<1) | (v>0)), or the offsets themselves (direct).  *bits_per_key (if
 * given) receives getBitsPerKey() (mphf.hpp:88-91) with the reference's byte
 * accounting. */
uint64_t orc_encode_mphf(const uint32_t* pilots, const uint32_t* offsets, uint32_t np, uint32_t B,
                         uint32_t P, const uint32_t* fulcs, uint64_t n, uint64_t seed, uint32_t encoding,
                         uint8_t* out, double* bits_per_key) {
  uint8_t* o = out;
  uint64_t total = 0, pilot_bytes = 0, off_bytes = 0;
  const uint32_t pid = encoding & 0xFF, oid = (encoding >> 8) & 0xFF;
  o = put64(o, 0x3248504d424f4850ull); /* "PHOBMPH2" */
  o = put64(o, n);
  o = put64(o, P);
  o = put64(o, np);
  o = put64(o, B);
  o = put64(o, seed);
  o = put64(o, encoding);
  o = put64(o, 1000000u);
  total += 64;
  for (int i = 0; i < ORC_FULCS; i += 2) o = put64(o, (uint64_t)fulcs[i] | ((uint64_t)fulcs[i + 1] << 32));
  total += ORC_FULCS * 4;
  const uint64_t nb = (uint64_t)np * B;
  uint64_t* col = (uint64_t*)malloc(((pid == 2 || pid == 3) ? nb + 1 : (uint64_t)np + 1) * 8);
  if (np && (pid == 2 || pid == 3)) { /* one sequence in build order p + b*np (mono_encoders.hpp:14-18) */
    o = put64(o, np);
    total += 8;
    for (uint64_t i = 0; i < nb; i++) col[i] = pilots[i];
    total += put_seq(&o, col, nb, pid - 2, &pilot_bytes);
  } else if (np) {
    /* columns; dual: the first floor(B * 0.5f) with the first base
     * (interleaved_encoder_dual.hpp:10-17), the split written first */
    const int dual = pid >= 4;
    const uint64_t b1 = dual ? (uint64_t)((float)B * 0.5f) : B;
    if (dual) {
      o = put64(o, b1);
      total += 8;
    }
    for (uint32_t j = 0; j < B; j++) {
      const uint32_t base = dual ? (j < b1 ? (pid - 4) >> 1 : (pid - 4) & 1) : pid;
      for (uint32_t p = 0; p < np; p++) col[p] = pilots[(uint64_t)j * np + p];
      total += put_seq(&o, col, np, base, &pilot_bytes);
    }
  }
  if (oid >= 2) { /* offsets as they are (direct_partition_offset_encoder.hpp:14-17) */
    for (uint32_t i = 0; i <= np; i++) col[i] = offsets[i];
    total += put_seq(&o, col, (uint64_t)np + 1, oid - 2, &off_bytes);
  } else {
    for (uint32_t i = 0; i <= np; i++) {
      int64_t v = (int64_t)offsets[i] - (int64_t)i * (int64_t)P;
      uint64_t a = (uint64_t)(v < 0 ? -v : v);
      col[i] = (a << 1) | (uint64_t)(v > 0);
    }
    o = put64(o, P);
    total += 8;
    total += put_seq(&o, col, (uint64_t)np + 1, oid, &off_bytes);
  }
  free(col);
  if (bits_per_key)
    *bits_per_key =
        n ? (double)(8 * pilot_bytes + 8 * off_bytes + 32ull * ORC_FULCS + 32ull * np) / (double)n : 0.0;
  return total;
}
