// search_kernel.hip -- per-partition bucket sort + pilot search (gfx950).
//
// One wave64 per partition; a workgroup holds kWavesPerGroup independent
// partitions and every wave synchronises only with itself.  Per partition:
//   1. bucket sizes from the partition's records (LDS atomics)
//      -- bucket_sizes.comp:21-35
//   2. processing order: size descending, bucket index descending
//      -- sort_sum.comp:156-199 (serial lane-0 loop in the reference; here
//      a ballot-ranked counting sort)
//   3. bucket start offsets inside the partition -- sort_sum.comp:203-225
//   4. keys written to the partition's slice of `sorted` in bucket order
//      -- redistribute_keys.comp:21-53
//   5. pilot search -- search.comp:31-175: for every bucket in order, the
//      smallest V = d + ps*pilot such that the positions
//      (hash(l1, hash(l2, pilot)) >> 1) mod ps + d (mod ps) are pairwise
//      distinct and free; they are then marked taken.
//
// Slot state lives in an extended bitmap E with E[i] = taken[i mod tm] for
// i < 2*tm, so the candidate offsets [D, D+w) of a key at initial position p
// are one unaligned w-bit read of E at p + D.  Every bit that decides an
// offset d < tm lies below 2*tm (p < tm); reads past it only meet offsets
// that the masks drop, so each slot is marked in two copies.  A trial scans all
// offsets bit-parallel: lane l owns offsets [32 l, 32 l + 32) of a pass of
// 2048, reads the 32-bit window of every key (consecutive words across the
// lanes, so no bank conflicts) and ORs them; the lowest free bit of the
// lowest lane is the smallest valid offset, which is the reference's answer.
// Buckets of at least 6 keys in a table at most half full first try offsets
// 0..63 with one 64-bit window per key and a DPP OR-reduction, which nearly
// always succeeds there.
//
// Buckets come in three classes, in processing order: above 64 keys (their
// chunks of 64 enter the collision bitmap one at a time, so most pilots are
// dropped before any scan), 2..64 keys (lane groups hash several pilots per
// round; pilots below 1024 run in a loop without the rare checks, whose
// accepted pilot marks E directly and learns from the returned bits whether
// two of its slots coincide), and one key (pilot 0 always; 64 slots hashed
// at once, then scanned one by one).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "device_hash.hpp"
#include "kernels.hpp"
#include "search_layout.hpp"
#include "xcd.hpp"

namespace phobic {
namespace {

constexpr uint32_t kMaxBucket = 256;  // sort_sum.comp BINS, search.comp MAX_BUCKET_SIZE

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Orders this wave's LDS / global accesses before and after the call for
// all lanes of the wave (the partition is owned by one wave).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Orders this wave's LDS accesses (a wave's LDS operations execute in
// order; this waits for them and stops the compiler from moving memory
// operations across).  Pending global loads are not waited for.
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Identical keys collide for every pilot; a bucket still colliding at this
// pilot is checked for them once (the outcome equals checking at pilot 0).
constexpr uint32_t kDupCheckPilot = 1024;

// records loaded per lane per step of the partition prologue
constexpr uint32_t kLoadBatch = 8;

// smallest bucket that tries offsets 0..63 first (while the table is at most
// half full); below it the full scan alone is cheaper
constexpr uint32_t kWin0MinKeys = 6;

// smallest table for which mod_magic's 24-bit product is exact (quotients
// below 2^22)
constexpr uint32_t kMagicModMin = 512;

template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, BOUND);
}

// OR over the 64 lanes (all lanes active), uniform result.
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  v |= dpp<0x111, 0xF, true>(v);   // row_shr:1
  v |= dpp<0x112, 0xF, true>(v);   // row_shr:2
  v |= dpp<0x114, 0xF, true>(v);   // row_shr:4
  v |= dpp<0x118, 0xF, true>(v);   // row_shr:8
  v |= dpp<0x142, 0xA, false>(v);  // row_bcast:15
  v |= dpp<0x143, 0xC, false>(v);  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// 64 bits of E starting at the packed slot x.
__device__ __forceinline__ uint64_t window64(const uint32_t* E, uint32_t x) {
  const uint32_t* ew = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(E) + (x >> 5));
  const uint32_t a = ew[0], b = ew[1], c = ew[2];
  return ((uint64_t)__builtin_amdgcn_alignbit(c, b, x) << 32) | __builtin_amdgcn_alignbit(b, a, x);
}

// Lemire's fastmod (M = floor((2^64-1)/d) + 1): a mod d for 32-bit a, d.
__device__ __forceinline__ uint32_t fastmod(uint32_t a, uint64_t M, uint32_t d) {
  return (uint32_t)__umul64hi(M * a, (uint64_t)d);
}

// a mod d for a < 2^31 and 512 <= d < 2^24 by a multiply with a magic
// number (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994, Theorem 4.2): with l = ceil(log2 d) and
// m = ceil(2^(31+l) / d), m*d - 2^(31+l) < d <= 2^l, so floor(a / d) =
// floor(m*a / 2^(31+l)) = mulhi(a, m) >> (l - 1) for every a < 2^31, and
// m < 2^32 (m = 2^31 when d is a power of two).  The quotient is below 2^22,
// so q*d is one 24x24-bit product.
__device__ __forceinline__ uint32_t mod_magic(uint32_t a, uint32_t m, uint32_t shift, uint32_t d) {
  const uint32_t q = __umulhi(a, m) >> shift;
  // the 24x24-bit product as one instruction (the compiler otherwise emits a
  // mask and a v_mul_lo_u32 in some loops)
  uint32_t qd;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(qd) : "s"(d), "v"(q));
  return a - qd;
}

__device__ __forceinline__ uint64_t low_mask(uint32_t bits) {  // bits >= 64 -> all ones
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}


// Packed slot q: (q >> 5) << 7 | (q & 31).  Shifted right by 5 it is the
// byte offset of q's bitmap word; v_alignbit reads only its low 5 bits.
__device__ __forceinline__ uint32_t pack_slot(uint32_t q) { return q + __umul24(q & ~31u, 3u); }
__device__ __forceinline__ uint32_t unpack_slot(uint32_t x) { return ((x >> 2) & ~31u) | (x & 31); }

// Mask of the offsets [off, off + 32) below lim.
__device__ __forceinline__ uint32_t below_mask(uint32_t off, uint32_t lim) {
  return off >= lim ? 0u : lim - off >= 32 ? ~0u : (1u << (lim - off)) - 1;
}

// ORs into acc[h] the 32-bit windows at o0 + 2048 h + 32 lane (eb = the
// lane's E word of offset o0) of the N slots held by lanes lane0 ..
// lane0 + N - 1 of pk (packed); the N reads are issued before the first use.
template <int H, uint32_t N>
__device__ __forceinline__ void or_windows(const char* eb, uint32_t pk, uint32_t lane0, uint32_t (&acc)[H]) {
  uint32_t x[N], lo[N][H], hi[N][H];
#pragma unroll
  for (uint32_t i = 0; i < N; i++) {
    x[i] = __builtin_amdgcn_readlane(pk, lane0 + i);
    const uint32_t* ew = reinterpret_cast<const uint32_t*>(eb + (x[i] >> 5));
#pragma unroll
    for (int h = 0; h < H; h++) {
      lo[i][h] = ew[64 * h];
      hi[i][h] = ew[64 * h + 1];
    }
  }
#pragma unroll
  for (uint32_t i = 0; i < N; i++)
#pragma unroll
    for (int h = 0; h < H; h++) acc[h] |= __builtin_amdgcn_alignbit(hi[i][h], lo[i][h], x[i]);
}

// Each even lane's packed slot with its odd neighbour's in the high half
// (DPP quad_perm [1,0,3,2]); packed slots of tables up to 4096 slots are
// below 2^14.
__device__ __forceinline__ uint32_t pair_pack(uint32_t pk) {
  // (shifted before the lane move, so that the move folds into the OR)
  return pk | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(pk << 16), 0xB1, 0xF, 0xF, false);
}

// ORs into acc[h] the windows of the 2 NP slots held by lanes lane0 ..
// lane0 + 2 NP - 1 (lane0 even) of a pair_pack register: one v_readlane
// fetches two slots.  The first slot's bit shift is the low 5 bits of the
// pair itself, the second's those of the pair >> 16.
template <int H, uint32_t NP, bool kSet = false>
__device__ __forceinline__ void or_pairs(const char* eb, uint32_t pk2, uint32_t lane0, uint32_t (&acc)[H]) {
  uint32_t x[NP], lo[NP][2][H], hi[NP][2][H];
#pragma unroll
  for (uint32_t i = 0; i < NP; i++) {
    x[i] = __builtin_amdgcn_readlane(pk2, lane0 + 2 * i);
    const uint32_t* ew0 = reinterpret_cast<const uint32_t*>(eb + ((x[i] & 0xFFFFu) >> 5));
    const uint32_t* ew1 = reinterpret_cast<const uint32_t*>(eb + (x[i] >> 21));
#pragma unroll
    for (int h = 0; h < H; h++) {
      lo[i][0][h] = ew0[64 * h];
      hi[i][0][h] = ew0[64 * h + 1];
      lo[i][1][h] = ew1[64 * h];
      hi[i][1][h] = ew1[64 * h + 1];
    }
  }
#pragma unroll
  for (uint32_t i = 0; i < NP; i++)
#pragma unroll
    for (int h = 0; h < H; h++) {
      const uint32_t w = __builtin_amdgcn_alignbit(hi[i][0][h], lo[i][0][h], x[i]) |
                         __builtin_amdgcn_alignbit(hi[i][1][h], lo[i][1][h], x[i] >> 16);
      acc[h] = (kSet && i == 0) ? w : acc[h] | w;  // kSet: the first pair sets acc
    }
}

// The windows of all s slots of a bucket: lanes [gbase, gbase + s) of pk[0]
// when C = 1 (4 or 2, then 2, then 1 at a time, so no slot is read twice), or
// lane k & 63 of pk[k >> 6] when C > 1 (2 at a time, the last slot repeated
// when one is left; OR is idempotent).
template <int C, int H>
__device__ __forceinline__ void or_bucket(const char* eb, const uint32_t (&pk)[C], uint32_t gbase, uint32_t s,
                                          uint32_t (&acc)[H]) {
  if constexpr (C == 1 && H < 4) {
    uint32_t k = 0;
    for (; k + 4 <= s; k += 4) or_windows<H, 4>(eb, pk[0], gbase + k, acc);
    if (s & 2) {
      or_windows<H, 2>(eb, pk[0], gbase + k, acc);
      k += 2;
    }
    if (s & 1) or_windows<H, 1>(eb, pk[0], gbase + k, acc);
  } else if constexpr (C == 1) {  // four sub-windows: 2 slots at a time keep the registers low
    uint32_t k = 0;
    for (; k + 2 <= s; k += 2) or_windows<H, 2>(eb, pk[0], gbase + k, acc);
    if (s & 1) or_windows<H, 1>(eb, pk[0], gbase + k, acc);
  } else {
#pragma unroll
    for (int j = 0; j < C; j++) {
      if (64 * j < s) {
        const uint32_t kl = min(s - 64 * j, 64u) - 1;
        for (uint32_t k = 0; k <= kl; k += 2) {
          uint32_t x[2], lo[2][H], hi[2][H];
#pragma unroll
          for (uint32_t i = 0; i < 2; i++) {
            x[i] = __builtin_amdgcn_readlane(pk[j], gbase + min(k + i, kl));
            const uint32_t* ew = reinterpret_cast<const uint32_t*>(eb + (x[i] >> 5));
#pragma unroll
            for (int h = 0; h < H; h++) {
              lo[i][h] = ew[64 * h];
              hi[i][h] = ew[64 * h + 1];
            }
          }
#pragma unroll
          for (uint32_t i = 0; i < 2; i++)
#pragma unroll
            for (int h = 0; h < H; h++) acc[h] |= __builtin_amdgcn_alignbit(hi[i][h], lo[i][h], x[i]);
        }
      }
    }
  }
}

// ORs into acc[h] the windows at offsets [0, 2048 H) of the s slots of lane
// group gbase (C = 1): one pass of a table of at most 2048 H slots.
template <int H>
__device__ __forceinline__ void scan_or(const uint32_t* Elane, uint32_t pk, uint32_t pk2, uint32_t gbase, uint32_t s,
                                        uint32_t (&acc)[H]) {
  const char* eb = reinterpret_cast<const char*>(Elane);
  if constexpr (H < 4) {  // two slots per lane read (gbase is a multiple of 4)
    // s >= 2: the first pair sets the accumulators; then the lane index
    // itself is the loop variable
    or_pairs<H, 1, true>(eb, pk2, gbase, acc);
    uint32_t rest = s - 2;
    // tested here in each scan (the compiler would keep hoisted tests as
    // lane masks)
    asm volatile("" : "+s"(rest));
    uint32_t l = gbase + 2;
    const uint32_t lend = l + (rest & ~3u);
    for (; l < lend; l += 4) or_pairs<H, 2>(eb, pk2, l, acc);
    if (rest & 2) {
      or_pairs<H, 1>(eb, pk2, l, acc);
      l += 2;
    }
    if (rest & 1) or_windows<H, 1>(eb, pk, l, acc);
  } else {
#pragma unroll
    for (int h = 0; h < H; h++) acc[h] = 0;
    const uint32_t pka[1] = {pk};
    or_bucket<1, H>(eb, pka, gbase, s, acc);
  }
}

// The smallest free offset of a pass.  fr[h]: this lane's free offsets of
// sub-window h (bit i: offset 2048 h + 32 lane + i); anyall: the ballot of
// the lanes with any, nonzero.
template <int H>
__device__ __forceinline__ uint32_t first_free(const uint32_t (&fr)[H], uint64_t anyall) {
  if constexpr (H == 1) {
    const uint32_t l0 = (uint32_t)__builtin_ctzll(anyall);
    return 32 * l0 + (uint32_t)__builtin_ctz(__builtin_amdgcn_readlane(fr[0], l0));
  } else {
    uint32_t d = ~0u;
#pragma unroll
    for (int h = H - 1; h >= 0; h--) {
      const uint64_t any = __ballot(fr[h] != 0);
      if (any) {
        const uint32_t l0 = (uint32_t)__builtin_ctzll(any);
        d = 2048 * h + 32 * l0 + (uint32_t)__builtin_ctz(__builtin_amdgcn_readlane(fr[h], l0));
      }
    }
    return d;
  }
}

// Smallest offset d < dlimit at which the s slots of the bucket (held in pk
// as or_bucket reads them) are all free, or false.  One pass covers
// 2048 * H offsets: lane l ORs the 32-bit windows of every slot at
// o0 + 2048 h + 32 l (consecutive words across the lanes); the lowest free
// bit of the lowest lane is the smallest offset.  `vfull` is the first
// pass's offset mask for dlimit = tm.
template <int C, int H>
__device__ __forceinline__ bool scan_offsets(const uint32_t* Elane, const uint32_t (&pk)[C], uint32_t gbase,
                                             uint32_t s, uint32_t dlimit, bool full, const uint32_t (&vfull)[H],
                                             uint32_t& dsel) {
  const uint32_t lane = __lane_id();
  if (full && dlimit <= 2048 * H) {  // one pass, offset masks known
    uint32_t acc[H];
#pragma unroll
    for (int h = 0; h < H; h++) acc[h] = 0;
    or_bucket<C, H>(reinterpret_cast<const char*>(Elane), pk, gbase, s, acc);
#pragma unroll
    for (int h = 0; h < H; h++) {
      const uint32_t fr = ~acc[h] & vfull[h];
      const uint64_t any = __ballot(fr != 0);
      if (any) {
        const uint32_t l0 = (uint32_t)__builtin_ctzll(any);
        dsel = 2048 * h + 32 * l0 + (uint32_t)__builtin_ctz(__builtin_amdgcn_readlane(fr, l0));
        return true;
      }
    }
    return false;
  }
  for (uint32_t o0 = 0; o0 < dlimit; o0 += 2048 * H) {
    uint32_t acc[H];
#pragma unroll
    for (int h = 0; h < H; h++) acc[h] = 0;
    or_bucket<C, H>(reinterpret_cast<const char*>(Elane + (o0 >> 5)), pk, gbase, s, acc);
#pragma unroll
    for (int h = 0; h < H; h++) {
      const uint32_t valid = (full && o0 == 0) ? vfull[h] : below_mask(o0 + 2048 * h + 32 * lane, dlimit);
      const uint32_t fr = ~acc[h] & valid;
      const uint64_t any = __ballot(fr != 0);
      if (any) {
        const uint32_t l0 = (uint32_t)__builtin_ctzll(any);
        dsel = o0 + 2048 * h + 32 * l0 + (uint32_t)__builtin_ctz(__builtin_amdgcn_readlane(fr, l0));
        return true;
      }
    }
  }
  return false;
}
}  // namespace

template <uint32_t kWavesPerGroup>
__global__ __launch_bounds__(64 * kWavesPerGroup) __attribute__((amdgpu_waves_per_eu(8))) void k_partition_search(
    SearchArgs a) {
  extern __shared__ __align__(16) uint32_t smem[];
  const uint32_t lane = __lane_id();
  // wave-uniform values are made visibly uniform so that they live in
  // SGPRs and loops over them stay scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t cap = a.cap, tcap = a.tcap;
  // B: buckets per partition (of this partition's set, below); a.B sizes the
  // LDS layout and is the largest of them
  uint32_t B = a.B;
  const uint32_t EW = e_words(tcap), CW = (tcap + 31) / 32 + 1;
  // the offsets of search_layout (layout_matches_kernel checks them on the
  // host before every launch)
  uint32_t* rt = smem + wave * lds_words(tcap, a.B);  // 1 word per rank
  uint32_t* cp = rt + a.B;  // counts, then pilots, per bucket index
  uint32_t* U = cp + a.B;
  uint32_t* hist = U;          // steps 1-2
  uint32_t* bpos = U;          // steps 3-4
  uint32_t* E = U;             // steps 5-6
  uint32_t* coll = U + EW;
  uint32_t* bcnt = cp;
  uint32_t* pil = cp;

  // consecutive partitions on one XCD: the interleaved layout p + b*np puts
  // 16 partitions' pilots in one 64-byte line
  const uint32_t p = xcd_block(blockIdx.x, gridDim.x) * kWavesPerGroup + wave;
  if (p >= a.np) return;
  const uint32_t ps = __builtin_amdgcn_readfirstlane(a.pcount[p]);
  const uint64_t base = (uint64_t)p * cap;
  // output: pilots[p + b*np] (one build) or the owning set's slice
  uint32_t* out_pilots = a.pilots;
  uint32_t out_p = p, out_np = a.np;
  if (a.sets) {
    uint32_t lo_s = 0, hi_s = a.sets;
    while (hi_s - lo_s > 1) {
      const uint32_t mid = (lo_s + hi_s) / 2;
      if (a.set_part_base[mid] <= p) lo_s = mid;
      else hi_s = mid;
    }
    // skip empty sets that share the same base
    while (lo_s + 1 < a.sets && a.set_part_base[lo_s + 1] <= p) lo_s++;
    const uint32_t pb = a.set_part_base[lo_s];
    if (a.set_B) {  // sets of several configurations
      B = __builtin_amdgcn_readfirstlane(a.set_B[lo_s]);
      out_pilots = a.pilots + a.set_pilot_base[lo_s];
    } else {
      out_pilots = a.pilots + (uint64_t)B * pb;
    }
    out_p = p - pb;
    out_np = a.set_part_base[lo_s + 1] - pb;
  }
  if (ps > cap) {  // the scatter stage already raised kErrPartitionOverflow
    for (uint32_t b = lane; b < B; b += 64) out_pilots[out_p + (uint64_t)b * out_np] = 0;
    return;
  }
  for (uint32_t i = lane; i < B; i += 64) bcnt[i] = 0;
  for (uint32_t i = lane; i < 256; i += 64) hist[i] = 0;
  lds_sync();

  // 1. bucket sizes (kLoadBatch loads in flight per lane)
  for (uint32_t i0 = 0; i0 < ps; i0 += 64 * kLoadBatch) {
    uint32_t m[kLoadBatch];
#pragma unroll
    for (uint32_t j = 0; j < kLoadBatch; j++) {
      const uint32_t i = i0 + lane + 64 * j;
      m[j] = i < ps ? a.rec_bucket[base + i] : 0xFFFFu;  // buckets are below B <= 65535
    }
#pragma unroll
    for (uint32_t j = 0; j < kLoadBatch; j++)
      if (m[j] != 0xFFFFu) atomicAdd(&bcnt[m[j]], 1u);
  }
  lds_sync();
  bool over = false;
  for (uint32_t b = lane; b < B; b += 64) {
    const uint32_t s = bcnt[b];
    if (s > kMaxBucket) over = true;
    else if (s) atomicAdd(&hist[256 - s], 1u);
  }
  if (__any(over)) {
    if (lane == 0) atomicOr(a.status, 1u << kErrBucketOverflow);
    for (uint32_t b = lane; b < B; b += 64) out_pilots[out_p + (uint64_t)b * out_np] = 0;
    return;
  }
  lds_sync();

  // 2a. bin starts: exclusive scan over t = 256 - size (largest first)
  uint32_t m, r_big, n_one;  // non-empty buckets; those with more than 64 keys; with one key
  {
    uint32_t v[4], loc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      v[j] = hist[4 * lane + j];
      loc += v[j];
    }
    const uint32_t incl = wave_incl_scan(loc);
    uint32_t run = incl - loc;
    m = __builtin_amdgcn_readlane(incl, 63);
    r_big = __builtin_amdgcn_readlane(run, (256 - 64) / 4);  // bins t < 192: sizes 65..256
    n_one = __builtin_amdgcn_readlane(v[3], 63);              // bin t = 255: size 1
#pragma unroll
    for (int j = 0; j < 4; j++) {
      hist[4 * lane + j] = run;
      run += v[j];
    }
  }
  lds_sync();

  // 2b. ranks: within one size, larger bucket index first
  for (uint32_t c = 0; c < B; c += 64) {
    const int idx = (int)B - 1 - (int)(c + lane);
    const uint32_t s = idx >= 0 ? bcnt[idx] : 0;
    bool act = s > 0;
    uint64_t am = __ballot(act);
    while (am) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(am);
      const uint32_t v = __builtin_amdgcn_readlane(s, leader);
      const bool mine = act && s == v;
      const uint64_t same = __ballot(mine);
      // Each size of the chunk has its own bin t, so no iteration reads what
      // another of this chunk writes: the chunk's writes are waited for
      // once, after it (a wave's LDS operations execute in order).
      const uint32_t t = 256 - v;
      const uint32_t bv = hist[t];
      if (mine) {
        rt[bv + __popcll(same & lanemask_lt())] = (uint32_t)idx;
        act = false;
      }
      if (lane == leader) hist[t] = bv + __popcll(same);
      am &= ~same;
    }
    lds_sync();
  }
  lds_sync();

  // Lane groups of a bucket of s <= 64 keys: 2^lg lanes per pilot (lg >= 2),
  // 64 / 2^lg pilots per hash round; one group (lg = 6) for s > 32 or with
  // single-pilot rounds.
  const bool multi = !(a.mode & kSearchSinglePilot);
  auto lg_of = [&](uint32_t sz) -> uint32_t {
    return multi && sz <= 32 ? (sz <= 4 ? 2u : 32u - __builtin_clz(sz - 1)) : 6u;
  };

  // the table size of step 5 (tm = ps unless alpha < 1)
  const uint32_t tm = __builtin_amdgcn_readfirstlane(table_size(ps, a.a_ppm));

  // 3. bucket start offsets: the rank table (bucket | size << 16 | lg << 25)
  // and each bucket's cursor.  Also r_w0: the buckets that try offsets 0..63
  // first (at least kWin0MinKeys keys, placed while the table is at most half
  // full) are the ranks below r_w0, since sizes fall and starts rise with
  // the rank.
  uint32_t r_w0 = 0;
  {
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < m; r0 += 64) {
      const uint32_t r = r0 + lane;
      const uint32_t b = r < m ? rt[r] : 0;
      const uint32_t s = r < m ? bcnt[b] : 0;
      const uint32_t incl = wave_incl_scan(s);
      const uint32_t st = carry + incl - s;
      if (r < m) {
        rt[r] = b | s << 16 | lg_of(s) << 25;
        bpos[b] = st;
      }
      r_w0 += __popcll(__ballot(r < m && s >= kWin0MinKeys && 2 * st < tm));
      carry += __builtin_amdgcn_readlane(incl, 63);
    }
  }
  lds_sync();

  // 4. keys in processing order into this partition's slice of `sorted`
  uint64_t* __restrict__ sorted = a.sorted + base;
  for (uint32_t i0 = 0; i0 < ps; i0 += 64 * kLoadBatch) {
    uint32_t m[kLoadBatch];
    uint64_t v[kLoadBatch];
#pragma unroll
    for (uint32_t j = 0; j < kLoadBatch; j++) {
      const uint32_t i = i0 + lane + 64 * j;
      m[j] = i < ps ? a.rec_bucket[base + i] : 0xFFFFu;
      v[j] = i < ps ? a.rec_lower[base + i] : 0;
    }
    // All of the batch's loads, waited for at once: the vector memory
    // counter also counts stores, so a wait for a later load would
    // otherwise hold each store until the one before it completed.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) (gfx9 encoding)
#pragma unroll
    for (uint32_t j = 0; j < kLoadBatch; j++) {
      if (m[j] != 0xFFFFu) {
        const uint32_t b = m[j];
        sorted[atomicAdd(&bpos[b], 1u)] = v[j];
      }
    }
  }
  wave_sync();  // the sorted slice is read back by other lanes
  // the union region becomes E and coll, the count array the pilots
  for (uint32_t i = lane; i < EW + CW; i += 64) U[i] = 0;
  for (uint32_t i = lane; i < B; i += 64) pil[i] = 0;
  lds_sync();

  // 5. pilot search over a table of tm slots
  const bool magic = tm >= kMagicModMin;
  // One 64-bit constant serves either remainder: fastmod's M for small
  // tables; for the others, mod_magic's m = ceil(2^(31+l) / tm) (low word)
  // and shift l - 1 (high word), l = ceil(log2 tm).  The f64 quotient is
  // within 2^-21 of the exact one, and a quotient that is not an integer
  // lies at least 1 / tm >= 2^-13 below the next integer, so its ceiling is
  // exact.
  uint64_t modc;
  if (magic) {
    const uint32_t l = 32 - __builtin_clz(tm - 1);
    modc = (uint64_t)(uint32_t)ceil(ldexp(1.0, 31 + (int)l) / (double)tm) | (uint64_t)(l - 1) << 32;
  } else {
    modc = ~0ull / (tm ? tm : 1) + 1;
  }
  // largest pilot whose offsets [0, tm) all keep V below 2^31
  const uint32_t pilot_full = __builtin_amdgcn_readfirstlane(tm ? (uint32_t)(0x80000000ull / tm) - 1 : 0);
  const uint32_t* Elane = E + lane;
  uint32_t err = 0;
  // packed slot of key kv for pilot pl: (hash(l1, hash(l2, pl)) >> 1) mod tm
  // (mc: std::integral_constant<bool, magic>, so that the pilot loops hold
  // no branch on the remainder's kind)
  auto slot_q = [&](uint64_t kv, uint32_t pl, auto mc) {  // unpacked
    uint32_t h = mix32((uint32_t)kv, mix32((uint32_t)(kv >> 32), pl));
    // (k >> 16) ^ k is one instruction; folded into the shift below it
    // would become two shifts and an xor
    asm volatile("" : "+v"(h));
    h >>= 1;
    if constexpr (decltype(mc)::value) return mod_magic(h, (uint32_t)modc, (uint32_t)(modc >> 32), tm);
    else return fastmod(h, modc, tm);
  };
  auto slot = [&](uint64_t kv, uint32_t pl, auto mc) { return pack_slot(slot_q(kv, pl, mc)); };

  struct Bucket {
    uint32_t b = 0, s = 0, lg = 6, st = 0;  // index, keys, lane-group log2, start in the slice (wave-uniform)
  };
  auto bucket_at = [&](uint32_t rank, uint32_t start) {
    const uint32_t e = __builtin_amdgcn_readfirstlane(rt[rank]);
    Bucket k;
    k.b = e & 0xFFFFu;
    k.s = (e >> 16) & 0x1FFu;
    k.lg = e >> 25;
    k.st = start;
    return k;
  };
  uint32_t placed = 0, r = 0;

  // Buckets r < r_end whose keys fit in C registers per lane (C chunks of
  // 64), scanned with H sub-windows per lane.  Buckets of at most 32 keys
  // (C = 1) hash G = 64 / 2^lg consecutive pilots per round: lane group g
  // of 2^lg >= 4 lanes holds the slots for the round's pilot g, its lanes past s
  // repeating the last key.
  // w0: std::integral_constant<int, W>; W = 1: every bucket of the range
  // tries offsets 0..63 first, W = 0: none does, W = 2: decided per bucket
  auto search_buckets = [&](auto chunks, auto subwin, uint32_t r_end, auto mc,
                            auto w0)__attribute__((always_inline)) {
    constexpr int W = decltype(w0)::value;
    constexpr int C = decltype(chunks)::value;
    constexpr int H = decltype(subwin)::value;
    constexpr bool kPad = C == 1;
    if (r >= r_end) return;
    // A bucket's keys in registers: with C > 1 key k in lane k & 63 of
    // chunk k >> 6; with C = 1 lane l holds key min(l mod 2^lg, s - 1), so
    // every lane group holds the bucket (lanes past s repeat its last key).
    auto load_keys = [&](const Bucket& k, uint32_t lg, uint64_t (&dst)[C]) {
      if constexpr (C == 1) {
        dst[0] = sorted[k.st + min(lane & ((1u << lg) - 1), k.s - 1)];
      } else {
#pragma unroll
        for (int j = 0; j < C; j++) {
          const uint32_t kk = lane + 64 * j;
          dst[j] = kk < k.s ? sorted[k.st + kk] : 0;
        }
      }
    };
    uint32_t vfull[H];  // first-pass offset masks for dlimit = tm
#pragma unroll
    for (int h = 0; h < H; h++) vfull[h] = below_mask(2048 * h + 32 * lane, tm);
    // the end of the lean loop's pilots (one pass must cover the table: not
    // in the narrow mode above 2048 slots)
    const uint32_t fast_end =
        __builtin_amdgcn_readfirstlane(tm <= 2048 * H ? min(kDupCheckPilot, pilot_full + 1) : 0u);
    Bucket cur = bucket_at(r, placed);
    uint32_t clg = C == 1 ? cur.lg : 6u;
    uint64_t key[C];  // the current bucket's keys, in the lane-group layout
    load_keys(cur, clg, key);
    // in before the loop, so that inside it only the next bucket's keys
    // are ever in flight (a wait for older loads then leaves them alone)
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) (gfx9 encoding)
    for (; r < r_end; r++) {
      // prefetch the next bucket's keys
      // (the last bucket reads its own entry again: no defaults to set)
      const Bucket nxt = bucket_at(min(r + 1, r_end - 1), placed + cur.s);
      const uint32_t nlg = C == 1 ? nxt.lg : 6u;
      uint64_t nkey[C];
      if (r + 1 < r_end) load_keys(nxt, nlg, nkey);

      const uint32_t s = cur.s, cst = cur.st, cb = cur.b;
      // the d < 64 window test pays off for large buckets in a sparse table
      // s >= kWin0MinKeys && 2 placed < tm, as the sign bit of integer
      // differences (no boolean, which would be kept as a lane mask)
      uint32_t win0 = W;
      if constexpr (W == 2) win0 = ((kWin0MinKeys - 1 - s) & (2 * placed - tm)) >> 31;
      const uint32_t lg = clg;
      const uint32_t G = 64u >> lg;
      const uint32_t g = lane >> lg, kk = lane & ((1u << lg) - 1);
      uint32_t pk[C];  // packed slots
      // pk[0] paired (pair_pack) and unpacked, in the lean loop (set by its
      // first hash round, pilot 0)
      uint32_t pk2, qs;
      uint32_t pilot = 0, dsel = 0;
      bool dupchk = false;
      uint32_t done = 0;  // 1: placed and marked by the lean loop (wave-uniform)
      if constexpr (C == 1) {
        // The common case, with less control flow: pilots whose offsets all
        // keep V below 2^31 and that are below the duplicate check.  The
        // bucket has at least two keys (one-key buckets are placed apart).
        // Marks pilot group gi's slots at offset d in every copy of E and
        // tells whether they are pairwise distinct.  The scan found them all
        // free, so a bit that is already set means two of the bucket's own
        // slots coincide: the marks are then taken back.
        auto mark_distinct = [&](uint32_t gi, uint32_t d) {
          const bool mine = g == gi && kk < s;
          uint32_t q = qs + d;
          if (q >= tm) q -= tm;
          const uint32_t q1 = q + tm;  // both below 2 tm
          const uint32_t m0 = 1u << (q & 31), m1 = 1u << (q1 & 31);
          // word indices kept as such, so that each address is one
          // shift-and-add
          uint32_t w0 = q >> 5, w1 = q1 >> 5;
          asm volatile("" : "+v"(w0), "+v"(w1));
          uint32_t old = 0;  // (tested after the branch: one compare gives the ballot)
          if (mine) {
            old = atomicOr(&E[w0], m0);
            atomicOr(&E[w1], m1);
          }
          lds_sync();
          if (__ballot((old & m0) != 0) == 0) return 1u;
          if (mine) {
            atomicAnd(&E[w0], ~m0);
            atomicAnd(&E[w1], ~m1);
          }
          lds_sync();
          return 0u;
        };
        // The pilot loops have one exit, at the bottom, so that the
        // compiler keeps no exit flags in them.
        if (win0) {
          // loop state as integers (boolean state is kept as lane masks)
          uint32_t more = pilot < fast_end ? 1u : 0u, okl = 0;
          while (more) {
            const uint32_t gi = pilot & (G - 1);
            if (gi == 0) {
              qs = slot_q(key[0], pilot + g, mc);
              pk[0] = pack_slot(qs);
              pk2 = pair_pack(pk[0]);
            }
            const uint64_t t0 = g == gi ? window64(E, pk[0]) : 0;
            const uint64_t free0 = ~(((uint64_t)wave_or((uint32_t)(t0 >> 32)) << 32) | wave_or((uint32_t)t0)) &
                                   low_mask(tm);
            uint32_t ok = 0;  // 1: placed and marked (wave-uniform)
            if (free0) {
              dsel = (uint32_t)__builtin_ctzll(free0);
              ok = mark_distinct(gi, dsel);
            } else {
              uint32_t acc[H], fr[H], frall = 0;
              scan_or<H>(Elane, pk[0], pk2, gi << lg, s, acc);
#pragma unroll
              for (int h = 0; h < H; h++) {
                fr[h] = ~acc[h] & vfull[h];
                frall |= fr[h];
              }
              const uint64_t anyall = __ballot(frall != 0);
              if (anyall) {
                dsel = first_free<H>(fr, anyall);
                ok = mark_distinct(gi, dsel);
              }
            }
            pilot += 1 - ok;
            okl = ok;
            more = pilot < fast_end ? ok ^ 1u : 0u;
            asm volatile("" : "+s"(more));  // an integer in an SGPR, not a lane mask
          }
          done = okl;
          asm volatile("" : "+s"(done));
        } else {
          // loop state as integers (boolean state is kept as lane masks)
          uint32_t more = pilot < fast_end ? 1u : 0u, okl = 0;
          while (more) {
            const uint32_t gi = pilot & (G - 1);
            if (gi == 0) {
              qs = slot_q(key[0], pilot + g, mc);
              pk[0] = pack_slot(qs);
              pk2 = pair_pack(pk[0]);
            }
            uint32_t acc[H], fr[H], frall = 0;
            scan_or<H>(Elane, pk[0], pk2, gi << lg, s, acc);
#pragma unroll
            for (int h = 0; h < H; h++) {
              fr[h] = ~acc[h] & vfull[h];
              frall |= fr[h];
            }
            // one vector compare decides the common case, no free offset at all
            const uint64_t anyall = __ballot(frall != 0);
            uint32_t ok = 0;  // 1: placed and marked (wave-uniform)
            if (anyall) {
              dsel = first_free<H>(fr, anyall);
              ok = mark_distinct(gi, dsel);
            }
            pilot += 1 - ok;
            okl = ok;
            more = pilot < fast_end ? ok ^ 1u : 0u;
            asm volatile("" : "+s"(more));  // an integer in an SGPR, not a lane mask
          }
          done = okl;
          asm volatile("" : "+s"(done));
        }
      }
      for (; !done; pilot++) {
        uint32_t dlimit = tm;  // d < dlimit keeps V = d + tm * pilot below 2^31
        if (pilot > pilot_full) {
          const uint64_t vbase = (uint64_t)tm * pilot;
          if (vbase > 0x7FFFFFFFull) {
            err = kErrPilotOverflow;
            break;
          }
          dlimit = (uint32_t)(0x80000000ull - vbase);
        }
        const uint32_t gi = pilot & (G - 1);
        bool coincide = false;  // two slots of the bucket coincide (wave-uniform)
        if constexpr (C > 1) {
          // Above 64 keys most pilots fail because two slots coincide: the
          // chunks of 64 keys enter the collision bitmap one at a time and
          // the pilot is dropped at the first coincidence, before its
          // offsets are scanned.
          uint32_t jn = 0;
#pragma unroll
          for (int j = 0; j < C; j++) {
            if (!coincide && 64 * j < s) {
              const uint32_t k = lane + 64 * j;
              bool cj = false;
              pk[j] = 0;
              if (k < s) {
                pk[j] = slot(key[j], pilot, mc);
                const uint32_t msk = 1u << (pk[j] & 31);
                cj = (atomicOr(&coll[pk[j] >> 7], msk) & msk) != 0;
              }
              lds_sync();
              coincide = __any(cj);
              jn = j + 1;
            }
          }
#pragma unroll
          for (int j = 0; j < C; j++)
            if ((uint32_t)j < jn && lane + 64 * j < s) coll[pk[j] >> 7] = 0;
          lds_sync();
        } else if (gi == 0) {  // slots for pilots pilot .. pilot + G - 1
          pk[0] = slot(key[0], pilot + g, mc);
        }
        if (!coincide) {
          bool found = false;
          if (win0) {  // offsets 0..63: one window per key of group gi, OR over keys
            uint64_t t0 = 0;
#pragma unroll
            for (int j = 0; j < C; j++) {
              const uint32_t k = kk + 64 * j;
              if (g == gi && 64 * j < s && (kPad || k < s)) t0 |= window64(E, pk[j]);
            }
            const uint64_t taken0 = ((uint64_t)wave_or((uint32_t)(t0 >> 32)) << 32) | wave_or((uint32_t)t0);
            const uint64_t free0 = ~taken0 & low_mask(dlimit);
            if (free0) {
              dsel = (uint32_t)__builtin_ctzll(free0);
              found = true;
            }
          }
          if (!found && !scan_offsets<C, H>(Elane, pk, gi << lg, s, dlimit, dlimit == tm, vfull, dsel)) continue;
          if constexpr (C == 1) {
            if (s > 1) {  // the slots must be pairwise distinct
              bool c = false;
              if (g == gi && kk < s) {
                const uint32_t msk = 1u << (pk[0] & 31);
                c = (atomicOr(&coll[pk[0] >> 7], msk) & msk) != 0;
              }
              lds_sync();
              if (g == gi && kk < s) coll[pk[0] >> 7] = 0;
              lds_sync();
              coincide = __any(c);
            }
          }
        }
        if (coincide) {
          if (pilot >= kDupCheckPilot && !dupchk) {
            dupchk = true;
            bool dup = false;
            for (uint32_t k2 = 0; k2 < s; k2++) {
              const uint64_t kv = sorted[cst + k2];
#pragma unroll
              for (int j = 0; j < C; j++) {
                const uint32_t k = lane + 64 * j;
                if (k < s && k != k2 && key[j] == kv) dup = true;
              }
            }
            if (__any(dup)) {
              err = kErrDuplicateKeys;
              break;
            }
          }
          continue;
        }
        break;
      }
      if (err) return;
      const uint32_t gsel = pilot & (G - 1), psel = pilot;
      // mark the bucket's slots in every copy of E (the lean loop has)
      if (!done) {
#pragma unroll
        for (int j = 0; j < C; j++) {
          const uint32_t k = kk + 64 * j;
          if (g == gsel && 64 * j < s && k < s) {
            uint32_t q = unpack_slot(pk[j]) + dsel;
            if (q >= tm) q -= tm;
            const uint32_t q1 = q + tm;  // both below 2 tm
            atomicOr(&E[q >> 5], 1u << (q & 31));
            atomicOr(&E[q1 >> 5], 1u << (q1 & 31));
          }
        }
      }
      // every lane stores the same value (no exec-mask switch)
      pil[cb] = dsel + tm * psel;
      placed += s;
      lds_sync();
      cur = nxt;
      clg = nlg;
#pragma unroll
      for (int j = 0; j < C; j++) key[j] = nkey[j];
    }
  };
  // processing order is size-descending: buckets above 64 keys come first.
  // They are few and meet a sparse table, so they use 2048-offset passes,
  // which keeps the kernel's register count (and so its occupancy) low.
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  // Buckets of one key come last.  Their pilot is always 0 (a free slot
  // exists while keys are left, and slots of one key cannot coincide), so
  // only the first free offset from the key's slot is searched: the slots of
  // 64 of them are hashed at once, then each is scanned and marked in turn.
  auto single_keys = [&](auto subwin, auto mc)__attribute__((always_inline)) {
    constexpr int H = decltype(subwin)::value;
    uint32_t vfull[H];
#pragma unroll
    for (int h = 0; h < H; h++) vfull[h] = below_mask(2048 * h + 32 * lane, tm);
    const uint32_t r0 = r, st0 = placed;  // one key each: rank r starts at st0 + r - r0
    for (; r < m; r += 64) {
      const uint32_t cnt = min(64u, m - r);
      uint32_t pk[1] = {0}, bi = 0;
      if (lane < cnt) {
        bi = rt[r + lane] & 0xFFFFu;
        pk[0] = slot(sorted[st0 + (r - r0) + lane], 0, mc);
      }
      for (uint32_t i = 0; i < cnt; i++) {
        uint32_t dsel = 0;
        scan_offsets<1, H>(Elane, pk, i, 1, tm, true, vfull, dsel);  // finds one: a slot is free
        uint32_t q = unpack_slot(__builtin_amdgcn_readlane(pk[0], i)) + dsel;
        if (q >= tm) q -= tm;
        const uint32_t qc = q + lane * tm;  // the two copies of slot q in E, one lane each
        if (lane < 2) atomicOr(&E[qc >> 5], 1u << (qc & 31));
        if (lane == 0) pil[__builtin_amdgcn_readlane(bi, i)] = dsel;
        lds_sync();
      }
    }
  };
  const uint32_t r_one = m - n_one;
  auto search_all = [&](auto mc)__attribute__((always_inline)) {
    using W0 = std::integral_constant<int, 0>;
    using W1 = std::integral_constant<int, 1>;
    using W2 = std::integral_constant<int, 2>;
    search_buckets(I4{}, I1{}, min(r_big, m), mc, W2{});
    if (err) return;
    const uint32_t r_w = __builtin_amdgcn_readfirstlane(min(r_w0, r_one));
    if (tm <= 2048 || (a.mode & kSearchNarrow)) {
      search_buckets(I1{}, I1{}, r_w, mc, W1{});
      if (!err) search_buckets(I1{}, I1{}, r_one, mc, W0{});
      if (!err) single_keys(I1{}, mc);
    } else if (tm <= 4096) {
      search_buckets(I1{}, I2{}, r_w, mc, W1{});
      if (!err) search_buckets(I1{}, I2{}, r_one, mc, W0{});
      if (!err) single_keys(I2{}, mc);
    } else {
      search_buckets(I1{}, I4{}, r_w, mc, W1{});
      if (!err) search_buckets(I1{}, I4{}, r_one, mc, W0{});
      if (!err) single_keys(I4{}, mc);
    }
  };
  if (magic) search_all(std::true_type{});
  else search_all(std::false_type{});
  if (err && lane == 0) atomicOr(a.status, 1u << err);
  lds_sync();

  // 6. alpha < 1: the k-th used slot of [ps, tm) stands for the k-th free
  // slot of [0, ps).  R[0..K) first receives the free slots in order, then
  // the high slots are filled in descending order (each reads R[rank] with
  // rank <= its own index before anything at or below it is overwritten).
  if (a.remap && !err && tm > ps) {
    uint32_t* R = a.remap + a.remap_offsets[p];
    const uint32_t W0 = (ps + 31) / 32, wpl = (W0 + 63) / 64;
    const uint32_t w_lo = min(W0, lane * wpl), w_hi = min(W0, (lane + 1) * wpl);
    auto free_word = [&](uint32_t w) {
      uint32_t fr = ~E[w];
      if (w == W0 - 1 && (ps & 31)) fr &= (1u << (ps & 31)) - 1;
      return fr;
    };
    uint32_t cnt = 0;
    for (uint32_t w = w_lo; w < w_hi; w++) cnt += __popc(free_word(w));
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t K = __builtin_amdgcn_readlane(incl, 63);
    uint32_t rank = incl - cnt;
    for (uint32_t w = w_lo; w < w_hi; w++) {
      uint32_t fr = free_word(w);
      while (fr) {
        const uint32_t bit = (uint32_t)__builtin_ctz(fr);
        fr &= fr - 1;
        R[rank++] = 32 * w + bit;
      }
    }
    wave_sync();
    const uint32_t H = tm - ps;
    uint32_t above = 0;
    for (int c = (int)((H + 63) / 64) - 1; c >= 0; c--) {
      const uint32_t j = 64 * (uint32_t)c + lane;
      const uint32_t q = ps + j;
      const bool used = j < H && ((E[q >> 5] >> (q & 31)) & 1u);
      const uint64_t bal = __ballot(used);
      const uint32_t cnt_c = __popcll(bal);
      const uint32_t r = K - above - cnt_c + __popcll(bal & lanemask_lt());
      const uint32_t src = used ? R[r] : 0u;
      wave_sync();
      if (j < H) R[j] = src;
      wave_sync();
      above += cnt_c;
    }
  }
  for (uint32_t b = lane; b < B; b += 64) out_pilots[out_p + (uint64_t)b * out_np] = pil[b];
}

size_t search_lds_bytes(uint32_t tcap, uint32_t B, uint32_t waves) { return (size_t)lds_words(tcap, B) * 4 * waves; }

hipError_t launch_partition_search(const SearchArgs& a, uint32_t waves, hipStream_t s) {
  if (a.np == 0) return hipSuccess;
  if (!layout_matches_kernel(a.tcap, a.B)) return hipErrorInvalidValue;
  if ((a.set_B != nullptr) != (a.set_pilot_base != nullptr) || (a.set_B && !a.sets)) return hipErrorInvalidValue;
  const uint32_t groups = (a.np + waves - 1) / waves;
  const size_t lds = search_lds_bytes(a.tcap, a.B, waves);
  switch (waves) {
    case 1: hipLaunchKernelGGL(k_partition_search<1>, dim3(groups), dim3(64), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_partition_search<2>, dim3(groups), dim3(128), lds, s, a); break;
    case 8: hipLaunchKernelGGL(k_partition_search<8>, dim3(groups), dim3(512), lds, s, a); break;
    default: hipLaunchKernelGGL(k_partition_search<4>, dim3(groups), dim3(256), lds, s, a); break;
  }
  return hipGetLastError();
}

}  // namespace phobic
