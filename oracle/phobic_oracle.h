/*
 * phobic_oracle.h -- CPU restatement of the PHOBIC construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * construction path is compared against.  Only tests, the smoke entry point
 * and the benchmark's CPU-baseline leg may load it; the product path never
 * calls it and has no CPU fallback.
 *
 * Every function names the reference file:line whose behaviour it restates
 * (paths relative to the stefanfred/PHOBIC-GPU tree).
 */
#ifndef PHOBIC_ORACLE_H
#define PHOBIC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical numbering to include/phobic_hip.h. */
#define ORC_OK 0
#define ORC_E_INVALID 1
#define ORC_E_PARTITION_OVERFLOW 5
#define ORC_E_BUCKET_OVERFLOW 6
#define ORC_E_DUPLICATE_KEYS 7
#define ORC_E_PILOT_OVERFLOW 8

#define ORC_FULCS 2048       /* shaders/constants.glsl:1 FULCS_INTER */
#define ORC_MAX_BUCKET 256   /* search.comp MAX_BUCKET_SIZE / sort_sum.comp BINS */

/* The reference's pre-hashed key (include/phobicGpu/hasher.hpp:7-20). */
typedef struct {
  uint32_t partitioner, bucketer, lower1, lower2;
} orc_key;

/* XXH3-128 (xxHash 0.8.x) of data[0..len) with a 64-bit seed.
 * Reference call sites: include/phobicGpu/hasher.hpp:34,43,52,61 (seed 0). */
void orc_xxh3_128(const void* data, size_t len, uint64_t seed, uint64_t* lo64, uint64_t* hi64);

/* hasher.hpp:36-37: Key = {lo32(low64), hi32(low64), lo32(high64), hi32(high64)}. */
void orc_hash_u64(const uint64_t* keys, uint64_t n, uint64_t seed, orc_key* out);
void orc_hash_bytes(const uint8_t* bytes, const uint64_t* offsets, uint64_t n, uint64_t seed,
                    orc_key* out);

/* mphf_config.h:26-55 + bucketers.h:28-39: B and the 2048-entry fulcrum table. */
uint32_t orc_bucket_count(double lambda, uint32_t partition_size);
void orc_fulcrums(double lambda, uint32_t partition_size, uint32_t* fulcs_out);

/* key_mapping.glsl:9-49 / mphf.hpp:42-49. */
uint32_t orc_hi32_mul(uint32_t a, uint32_t b);
uint32_t orc_partition_of(uint32_t partitioner, uint32_t np);
uint32_t orc_bucket_of(uint32_t bucketer, const uint32_t* fulcs);

/* hashing.glsl:11-17 / mphf.hpp:24-30 */
uint32_t orc_mix32(uint32_t key, uint32_t pilot);

/* Full build from pre-hashed keys (the reference's nohash path,
 * mphf_builder.h:201-216 with noHash()).  Outputs:
 *   pilots[np*B]   interleaved: pilots[p + b*np]  (search.comp:168-174)
 *   offsets[np+1]  exclusive partition offsets      (mphf_builder.h:254-257)
 * Optional intermediates (may be NULL):
 *   bucket_sizes[np*B]  key count of every bucket   (bucket_sizes.comp:21-35)
 *   order[np*B]         bucket indices in processing order per partition,
 *                       padded with 0xFFFFFFFF       (sort_sum.comp:156-199)
 * Returns ORC_OK or an error status. */
int orc_build_keys(const orc_key* keys, uint64_t n, double lambda, uint32_t partition_size,
                   uint32_t* pilots, uint32_t* offsets, uint32_t* bucket_sizes, uint32_t* order);

/* Load factor extension (no reference counterpart): alpha = a_ppm / 1e6,
 * table size m = ceil(ps * 1e6 / a_ppm); used slots of [ps, m) remapped in
 * order onto the free slots of [0, ps).  a_ppm = 1e6 is the reference. */
uint32_t orc_table_size(uint32_t ps, uint32_t a_ppm);
uint64_t orc_remap_capacity(uint64_t n, uint32_t np, uint32_t a_ppm);
int orc_build_keys_ex(const orc_key* keys, uint64_t n, double lambda, uint32_t partition_size, uint32_t a_ppm,
                      uint32_t* pilots, uint32_t* offsets, uint32_t* remap, uint32_t* remap_offsets,
                      uint32_t* bucket_sizes, uint32_t* order);
/* The same build with a custom Bucketer's B and 2048-entry fulcrum table
 * (mphf_config.h:34-42); every entry must be at most B << 16. */
int orc_build_keys_table(const orc_key* keys, uint64_t n, uint32_t partition_size, uint32_t B,
                         const uint32_t* fulcs, uint32_t a_ppm, uint32_t* pilots, uint32_t* offsets,
                         uint32_t* remap, uint32_t* remap_offsets);
void orc_query_keys_ex(const orc_key* keys, uint64_t n, const uint32_t* pilots, const uint32_t* offsets,
                       uint32_t np, uint32_t B, const uint32_t* fulcs, uint32_t a_ppm, const uint32_t* remap,
                       const uint32_t* remap_offsets, uint32_t* out);

/* Partitions of the key stream splitmix64(kseed) outputs 1..n (hashed
 * with XXH3 seed hseed, hasher.hpp:42-48), computed without holding all keys.
 * orc_scan_splitmix: per-partition key counts (np) and, if max_bucket is not
 * NULL, the largest bucket of each partition (saturating at 65535).
 * orc_build_splitmix_parts: builds only the listed partitions, each on its
 * own exactly as orc_build_keys_ex does: pilots_sel[j*B + b], counts_sel[j],
 * and for a_ppm < 1e6 the concatenated remap entries of the listed
 * partitions in list order (at most remap_sel_cap). */
int orc_scan_splitmix(uint64_t kseed, uint64_t n, uint64_t hseed, double lambda, uint32_t partition_size,
                      uint32_t* counts, uint16_t* max_bucket);
int orc_build_splitmix_parts(uint64_t kseed, uint64_t n, uint64_t hseed, double lambda, uint32_t partition_size,
                             uint32_t a_ppm, const uint32_t* parts, uint32_t nparts, uint32_t* pilots_sel,
                             uint32_t* counts_sel, uint32_t* remap_sel, uint64_t remap_sel_cap);

/* Sizes for a build: np = ceil(n/P), B = ceil(P/lambda). */
void orc_sizes(uint64_t n, double lambda, uint32_t partition_size, uint32_t* np, uint32_t* B);

/* Evaluate the MPHF defined by (pilots, offsets) for pre-hashed keys
 * (mphf.hpp:32-40,77-86).  out[i] = position of keys[i]. */
void orc_query_keys(const orc_key* keys, uint64_t n, const uint32_t* pilots,
                    const uint32_t* offsets, uint32_t np, uint32_t B, const uint32_t* fulcs,
                    uint32_t* out);

/* compact_vector (include/encoders/base/compact_vector.hpp:194-336):
 * width = max==0 ? 1 : bit length of max; words = ceil(n*w/64) + 1 (guard).
 * Returns the number of 64-bit words written (callers size words[] with
 * orc_compact_words).  *width_out receives the field width. */
uint64_t orc_compact_words(uint64_t n, uint64_t width);
uint64_t orc_compact_width(const uint64_t* values, uint64_t n);
uint64_t orc_compact_encode(const uint64_t* values, uint64_t n, uint64_t* words, uint64_t* width_out);
uint64_t orc_compact_access(const uint64_t* words, uint64_t width, uint64_t i);

/* Serialised MPHF bytes ("PHOBMPH2").  encoding: low byte the pilot layout
 * (0 compact columns = FastQueryMphf, 1 Rice columns = SmallSpaceMphf, 2/3 one
 * compact/Rice sequence, 4 + 2*first + second dual columns split at
 * floor(B * 0.5f), so 6 = Rice then compact), next byte the offsets (0/1
 * differences to i*P compact/Rice, 2/3 direct compact/Rice).  Returns the
 * byte count;
 * when out is NULL only the size is computed.  *bits_per_key, if non-NULL,
 * receives the bits-per-key figure. */
uint64_t orc_encode_mphf(const uint32_t* pilots, const uint32_t* offsets, uint32_t np, uint32_t B,
                         uint32_t P, const uint32_t* fulcs, uint64_t n, uint64_t seed, uint32_t encoding,
                         uint8_t* out, double* bits_per_key);
/* Rice parameter of a value sequence (rice_sequence.hpp:15-21). */
uint64_t orc_rice_param(const uint64_t* v, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
