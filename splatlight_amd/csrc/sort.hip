// Stable LSD radix sort of (u32 key, u32 value) pairs, onesweep form, for gfx950 (wave64).
//
// Used twice per frame: the depth order (16-bit keys of every splat slot, values = slot index; the result is
// the reference's stable order, shaders/radix.wgsl:1-139) and the tile order (tile ids of the (splat, tile)
// pairs emitted in depth order; a stable sort keeps every tile's list in depth order).
//
// One read of the keys builds every pass's digit histogram (sl_radix_hist + sl_radix_bases). Each pass is then one
// kernel: persistent workgroups take 2048- or 4096-key tiles in ticket order; inside a tile each wave ranks its
// elements by one ballot per digit bit (peers with the same digit) and a per-wave LDS digit counter, so the rank
// follows element order; the tile's digit counts are published with a decoupled look-back
// (aggregate / inclusive flags packed with the count in one word) and the tile is reordered by digit in LDS
// so the scatter writes runs of consecutive addresses.
#include "sl_device.h"

#include <type_traits>

#define RS_THREADS 256
#define RS_WAVES (RS_THREADS / SL_WAVE)
#define RS_TILE_MIN (RS_THREADS * 8)   // the smallest onesweep tile (8 keys per lane): sizes the look-back state
#define RS_FLAG_AGG (1u << 30)
#define RS_FLAG_INC (2u << 30)
#define RS_FLAG_MASK (3u << 30)
#define RS_VALUE_MASK ((1u << 30) - 1u)

namespace {

__device__ __forceinline__ uint32_t load_n(const uint32_t *n_ptr, uint32_t n_static) {
  return n_ptr ? __hip_atomic_load(n_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : n_static;
}

// exclusive scan of one value per thread over a 256-thread block; returns the block total in *total
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *s_tmp, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63u) s_tmp[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < RS_WAVES; k++) {
    const uint32_t t = s_tmp[k];
    before += (uint32_t)k < w ? t : 0u;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

}  // namespace

// Digit geometry of a BITS-bit pass: DIG digits; each of the 256 threads owns PER consecutive digits in the per-digit
// steps (BITS <= 8: one digit, threads past DIG idle).
template <uint32_t BITS>
struct RsDigits {
  static constexpr uint32_t DIG = 1u << BITS;
  static constexpr uint32_t PER = DIG > RS_THREADS ? DIG / RS_THREADS : 1u;
  static_assert(BITS >= 4u && BITS <= 11u, "4- to 11-bit digits");
};

// Digit histograms of every pass (at most 3) in one read: hist[p * DIG + digit], digits at shift0 + BITS p.
template <uint32_t BITS>
__global__ void __launch_bounds__(RS_THREADS)
sl_radix_hist(const uint32_t *__restrict__ keys, const uint32_t *n_ptr, uint32_t n_static, uint32_t shift0,
              uint32_t npasses, uint32_t *__restrict__ hist) {
  constexpr uint32_t DIG = RsDigits<BITS>::DIG;
  __shared__ uint32_t s_hist[3 * DIG];
  for (uint32_t k = threadIdx.x; k < 3u * DIG; k += RS_THREADS) s_hist[k] = 0u;
  __syncthreads();
  const uint32_t n = load_n(n_ptr, n_static);
  for (uint32_t i = blockIdx.x * RS_THREADS + threadIdx.x; i < n; i += gridDim.x * RS_THREADS) {
    const uint32_t key = keys[i];
    for (uint32_t p = 0; p < npasses; p++) atomicAdd(&s_hist[p * DIG + ((key >> (shift0 + BITS * p)) & (DIG - 1u))], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < npasses * DIG; k += RS_THREADS)
    if (s_hist[k]) atomicAdd(&hist[k], s_hist[k]);
}

// Exclusive digit offsets per pass, in place: one workgroup per pass.
template <uint32_t BITS>
__global__ void __launch_bounds__(RS_THREADS) sl_radix_bases(uint32_t *__restrict__ hist) {
  constexpr uint32_t DIG = RsDigits<BITS>::DIG, PER = RsDigits<BITS>::PER;
  __shared__ uint32_t s_tmp[RS_WAVES];
  uint32_t *h = hist + blockIdx.x * DIG;
  uint32_t v[PER], sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < PER; j++) {
    const uint32_t d = threadIdx.x * PER + j;
    v[j] = d < DIG ? h[d] : 0u;
    sum += v[j];
  }
  uint32_t total;
  uint32_t ex = block_excl_scan256(sum, s_tmp, &total);
#pragma unroll
  for (uint32_t j = 0; j < PER; j++) {
    const uint32_t d = threadIdx.x * PER + j;
    if (d < DIG) h[d] = ex;
    ex += v[j];
  }
}

// One stable pass on digit (key >> shift) & (DIG - 1). vals_in == nullptr: values are the element index.
// keys_out == nullptr: keys are not written (the last pass when only the order is wanted).
// status: >= ceil(n / tile) * DIG words, zero on entry; ticket: one word, zero on entry.
// RS_ITEMS keys per lane (8 or 16, tuned): a tile of 256 * RS_ITEMS keys. BITS: the digit width (8, or 11 so that a
// frame's tile ids sort in one pass); the wave digit counters are 16-bit past 256 digits (a wave ranks at most 1,024
// keys), which keeps the 11-bit pass at 64 KiB of LDS.
template <uint32_t RS_ITEMS, uint32_t BITS>
__global__ void __launch_bounds__(RS_THREADS)
sl_radix_pass(const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
              uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, const uint32_t *n_ptr,
              uint32_t n_static, uint32_t shift, const uint32_t *__restrict__ bases, uint32_t *status,
              uint32_t *ticket) {
  constexpr uint32_t RS_TILE = RS_THREADS * RS_ITEMS, RS_WAVE_SPAN = SL_WAVE * RS_ITEMS;
  constexpr uint32_t DIG = RsDigits<BITS>::DIG, PER = RsDigits<BITS>::PER;
  using WCount = typename std::conditional<(DIG > 256u), uint16_t, uint32_t>::type;
  __shared__ uint32_t s_keys[RS_TILE];
  __shared__ uint32_t s_vals[RS_TILE];
  __shared__ WCount s_whist[RS_WAVES * DIG];
  __shared__ uint32_t s_dexcl[DIG];
  __shared__ uint32_t s_gbase[DIG];
  __shared__ uint32_t s_tmp[RS_WAVES];
  __shared__ uint32_t s_tile;

  const uint32_t n = load_n(n_ptr, n_static);
  const uint32_t num_tiles = (n + RS_TILE - 1u) / RS_TILE;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  uint32_t digit_base[PER];
#pragma unroll
  for (uint32_t j = 0; j < PER; j++) {
    const uint32_t d = tid * PER + j;
    digit_base[j] = d < DIG ? bases[d] : 0u;
  }

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t k = tid; k < RS_WAVES * DIG; k += RS_THREADS) s_whist[k] = 0u;
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= num_tiles) break;
    const uint32_t tile_base = t * RS_TILE;
    const uint32_t tile_n = min(RS_TILE, n - tile_base);
    const uint32_t wbase = tile_base + w * RS_WAVE_SPAN;

    uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
      const uint32_t idx = wbase + (uint32_t)k * SL_WAVE + lane;
      const bool v = idx < n;
      key[k] = v ? keys_in[idx] : 0u;
      val[k] = v ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
    // rank within the wave, element order (k, lane): peers with the same digit by one ballot per digit bit
    WCount *whist = s_whist + w * DIG;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
      const uint32_t idx = wbase + (uint32_t)k * SL_WAVE + lane;
      const bool v = idx < n;
      const uint32_t d = (key[k] >> shift) & (DIG - 1u);
      uint64_t peers = __ballot(v);
#pragma unroll
      for (uint32_t b = 0; b < BITS; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      uint32_t r = 0;
      if (v) {
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        const uint32_t old = whist[d];
        r = old + below;
        if (below == 0u) whist[d] = (WCount)(old + (uint32_t)__popcll(peers));
      }
      rank[k] = r;
    }
    __syncthreads();
    // each owned digit: per-wave exclusive offsets and the tile's count, then its decoupled look-back
    uint32_t cnt[PER], excl[PER], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
      const uint32_t d = tid * PER + j;
      uint32_t c = 0;
      if (d < DIG) {
#pragma unroll
        for (int k = 0; k < RS_WAVES; k++) {
          const uint32_t x = s_whist[k * DIG + d];
          s_whist[k * DIG + d] = (WCount)c;
          c += x;
        }
      }
      cnt[j] = c;
      sum += c;
    }
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
      const uint32_t d = tid * PER + j;
      excl[j] = 0;
      if (d >= DIG) continue;
      uint32_t *st = status + (size_t)t * DIG + d;
      if (t == 0) {
        __hip_atomic_store(st, RS_FLAG_INC | cnt[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(st, RS_FLAG_AGG | cnt[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int32_t jj = (int32_t)t - 1;
        uint32_t e = 0;
        for (;;) {
          const uint32_t s = __hip_atomic_load(status + (size_t)jj * DIG + d, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
          const uint32_t fl = s & RS_FLAG_MASK;
          if (fl == 0u) { __builtin_amdgcn_s_sleep(1); continue; }
          e += s & RS_VALUE_MASK;
          if (fl == RS_FLAG_INC) break;
          --jj;
        }
        __hip_atomic_store(st, RS_FLAG_INC | (e + cnt[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        excl[j] = e;
      }
    }
    uint32_t total;
    uint32_t run = block_excl_scan256(sum, s_tmp, &total);
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
      const uint32_t d = tid * PER + j;
      if (d < DIG) {
        s_gbase[d] = digit_base[j] + excl[j];
        s_dexcl[d] = run;
      }
      run += cnt[j];
    }
    __syncthreads();
    // reorder the tile by digit in LDS
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
      const uint32_t idx = wbase + (uint32_t)k * SL_WAVE + lane;
      if (idx < n) {
        const uint32_t d = (key[k] >> shift) & (DIG - 1u);
        const uint32_t pos = s_dexcl[d] + (uint32_t)s_whist[w * DIG + d] + rank[k];
        s_keys[pos] = key[k];
        s_vals[pos] = val[k];
      }
    }
    __syncthreads();
    for (uint32_t j = tid; j < tile_n; j += RS_THREADS) {
      const uint32_t kk = s_keys[j];
      const uint32_t d = (kk >> shift) & (DIG - 1u);
      const uint32_t dst = s_gbase[d] + (j - s_dexcl[d]);
      if (keys_out) keys_out[dst] = kk;
      vals_out[dst] = s_vals[j];
    }
    __syncthreads();
  }
}

// Zero the look-back words a tile pass over the device's pair count uses: ceil(pairs / tile_keys) onesweep tiles of
// `digits` words each, at most cap words (a kernel reading the count, so the clear covers what the frame fills).
extern "C" __global__ void __launch_bounds__(256)
sl_fill_pairs_status(uint32_t *__restrict__ p, const uint32_t *pairs, uint32_t tile_keys, uint32_t digits,
                     uint32_t cap) {
  const uint32_t n = __hip_atomic_load(pairs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t words = (uint64_t)((n + tile_keys - 1u) / tile_keys) * digits;
  const uint32_t m = (uint32_t)(words < cap ? words : cap);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < m; i += gridDim.x * 256u) p[i] = 0u;
}

// Fill n words with a value (clears look-back state and counters; a kernel rather than a memset so the frame
// can be captured in a graph).
extern "C" __global__ void __launch_bounds__(256)
sl_fill_u32(uint32_t *__restrict__ p, uint32_t value, uint32_t n) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) p[i] = value;
}

// The same fill with words i0 and i1 of p set to v0 and v1 instead, and the word q (outside p) to qv: one launch for
// a pass's cleared scratch and its set-up words.
extern "C" __global__ void __launch_bounds__(256)
sl_fill_u32_set(uint32_t *__restrict__ p, uint32_t value, uint32_t n, uint32_t i0, uint32_t v0, uint32_t i1,
                uint32_t v1, uint32_t *__restrict__ q, uint32_t qv) {
  if (blockIdx.x == 0u && threadIdx.x == 0u) *q = qv;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u)
    p[i] = i == i0 ? v0 : (i == i1 ? v1 : value);
}

// ---- counting scatter (shaders/scatter.wgsl:9-18, the reference's ?sort=counting) ----------------------------------
// The depth order by a 65,536-bucket count of the 16-bit keys (project.wgsl:369), an exclusive scan of the counts
// (scan.wgsl:16-43) and one atomic placement per slot. Unlike the radix sort, equal keys land in whatever order their
// atomics arrive. Bucket 65,536 takes the invalid slots, so they still fill the order's tail after the visible ones.
#define CS_BUCKETS 65537u

extern "C" __global__ void __launch_bounds__(256)
sl_count_hist(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ hist) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  atomicAdd(&hist[k == 0xFFFFFFFFu ? 65536u : (k & 0xFFFFu)], 1u);
}

// one workgroup: each thread takes 65 consecutive buckets (1024 x 65 >= 65,537), in place
extern "C" __global__ void __launch_bounds__(1024)
sl_count_scan(uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_w[1024 / SL_WAVE];
  constexpr uint32_t PER = 65u;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t base = t * PER;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < PER; k++)
    if (base + k < CS_BUCKETS) sum += hist[base + k];
  uint32_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63u) s_w[w] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (uint32_t k = 0; k < w; k++) run += s_w[k];
  for (uint32_t k = 0; k < PER; k++) {
    if (base + k >= CS_BUCKETS) break;
    const uint32_t h = hist[base + k];
    hist[base + k] = run;
    run += h;
  }
}

extern "C" __global__ void __launch_bounds__(256)
sl_count_scatter(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *__restrict__ offsets,
                 uint32_t *__restrict__ order) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  order[atomicAdd(&offsets[k == 0xFFFFFFFFu ? 65536u : (k & 0xFFFFu)], 1u)] = i;
}
