// (splat, tile) pair emission and per-tile ranges.
//
// A pair is a visible splat and a 16x16 tile whose pixel-centre range meets the splat's extent box
// (shaders/project.wgsl:314). Pairs are emitted in the depth order produced by the depth sort, so after a
// stable sort by tile id every tile's list is that tile's slice of the reference's global back-to-front order.
//
// Three launches over 1024-rank chunks, none of which waits on another workgroup: sl_pairs_count writes each chunk's
// pair count, sl_pairs_scan (one workgroup) turns the counts into the chunks' first pair indices and the pair total,
// sl_pairs_emit scans its chunk's counts again and emits the chunk's pairs cooperatively in windows of PE_WIN: every
// rank whose first pair falls in the window marks that position with its rank, an inclusive max-scan over the
// window gives every position its owner, and lane j writes pair j, so the key and value stores are contiguous. (A
// single pass with chained chunk offsets spent a third of its time handing out tickets and waiting on its
// predecessors.) The tile-id digit histograms for the tile sort are accumulated on the way.
#include "sl_device.h"

#define PE_THREADS 512
#define PE_RPT 2
#define PE_CHUNK (PE_THREADS * PE_RPT)
#define PE_WIN 2048   // pairs per emission window

// The position of the k-th set bit (from 0) of m, which has more than k set bits.
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t k) {
  uint32_t pos = 0, c = (uint32_t)__popcll(m & 0xffffffffull);
  if (k >= c) { k -= c; m >>= 32; pos = 32; }
  uint32_t x = (uint32_t)m;
  c = (uint32_t)__popc(x & 0xffffu); if (k >= c) { k -= c; x >>= 16; pos += 16; }
  c = (uint32_t)__popc(x & 0xffu); if (k >= c) { k -= c; x >>= 8; pos += 8; }
  c = (uint32_t)__popc(x & 0xfu); if (k >= c) { k -= c; x >>= 4; pos += 4; }
  c = (uint32_t)__popc(x & 0x3u); if (k >= c) { k -= c; x >>= 2; pos += 2; }
  return pos + (k >= (x & 1u) ? 1u : 0u);
}

// The pair count of rank r (< ranks): the tiles of its slot's box, or (MASKED) the box's tiles set in its mask.
// order == nullptr: rank r is slot r.
template <bool MASKED>
__device__ __forceinline__ uint32_t pair_count(const uint32_t *__restrict__ order, const uint2 *__restrict__ tbox,
                                               const uint64_t *__restrict__ tmask, uint32_t r, uint32_t ranks,
                                               uint32_t &g, uint2 &b, uint64_t &m) {
  g = 0;
  b = make_uint2(0xffffffffu, 0u);
  m = 0;
  if (r >= ranks) return 0u;
  g = order ? order[r] : r;
  b = tbox[g];
  if (b.x == 0xffffffffu) return 0u;
  if (MASKED) {
    m = tmask[g];
    return (uint32_t)__popcll(m);
  }
  return ((b.y & 0xffffu) - (b.x & 0xffffu) + 1u) * ((b.y >> 16) - (b.x >> 16) + 1u);
}

// Exclusive scan of one value per thread over the workgroup (s_wsum: one word per wave); *total the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_wsum, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63u) s_wsum[w] = x;
  __syncthreads();
  uint32_t before = 0, t = 0;
#pragma unroll
  for (int k = 0; k < PE_THREADS / SL_WAVE; k++) {
    before += (uint32_t)k < w ? s_wsum[k] : 0u;
    t += s_wsum[k];
  }
  *total = t;
  return before + x - v;
}

// Chunk c's pair count into chunk_pairs[c] (the visible ranks are counters->visible; chunks past them exit).
template <bool MASKED>
__global__ void __launch_bounds__(PE_THREADS)
sl_pairs_count(const uint32_t *__restrict__ order, const uint2 *__restrict__ tbox, const uint64_t *__restrict__ tmask,
               const FrameCounters *counters, uint64_t *__restrict__ chunk_pairs) {
  __shared__ uint32_t s_wsum[PE_THREADS / SL_WAVE];
  const uint32_t V = __hip_atomic_load(&counters->visible, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t r0 = blockIdx.x * PE_CHUNK;
  if (r0 >= V) return;
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < PE_RPT; k++) {
    uint32_t g;
    uint2 b;
    uint64_t m;
    sum += pair_count<MASKED>(order, tbox, tmask, r0 + threadIdx.x * PE_RPT + k, V, g, b, m);
  }
  uint32_t total;
  (void)block_excl_scan(sum, s_wsum, &total);
  if (threadIdx.x == 0) chunk_pairs[blockIdx.x] = total;
}

// One 256-lane workgroup (small enough to start beside long-running kernels): the chunks' pair counts (chunk_pairs, in place) become their first pair indices, and
// chunk_pairs[chunks] the pair total; the total capped at the capacity and what overflowed go to the counters.
extern "C" __global__ void __launch_bounds__(256)
sl_pairs_scan(FrameCounters *counters, uint64_t *__restrict__ chunk_pairs, uint32_t capacity) {
  __shared__ uint64_t s_w[256 / SL_WAVE];
  __shared__ uint64_t s_carry;
  const uint32_t V = __hip_atomic_load(&counters->visible, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t chunks = (V + PE_CHUNK - 1u) / PE_CHUNK, lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < chunks; c0 += 256u) {
    const uint32_t c = c0 + threadIdx.x;
    const uint64_t v = c < chunks ? chunk_pairs[c] : 0ull;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) s_w[w] = x;
    __syncthreads();
    uint64_t before = s_carry, tot = 0;
#pragma unroll
    for (int k = 0; k < 256 / SL_WAVE; k++) {
      before += (uint32_t)k < w ? s_w[k] : 0ull;
      tot += s_w[k];
    }
    if (c < chunks) chunk_pairs[c] = before + x - v;
    __syncthreads();                                   // s_w and s_carry read before they change
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint64_t all = s_carry;
    chunk_pairs[chunks] = all;
    const uint32_t kept = all > capacity ? capacity : (uint32_t)all;
    __hip_atomic_store(&counters->pairs, kept, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&counters->overflow, (uint32_t)(all - kept), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Chunk c's pairs from its first pair index chunk_first[c]. Batches: a rank's slot belongs to the view v with
// slot_base[v] <= slot < slot_base[v + 1]; its pairs' tile ids are global, tile_base[v] + ty * tiles_x[v] + tx.
// tile_passes (1 to 3) digit histograms of BITS-bit digits (the tile sort's digit width) are built.
// MASKED (the opacity-map splats, boxes of at most 64 tiles): only the box's tiles whose bit (row * box width +
// column) is set in tmask[slot] get a pair; otherwise every tile of the box does.
template <bool MASKED, uint32_t BITS>
__global__ void __launch_bounds__(PE_THREADS)
sl_pairs_emit(const uint32_t *__restrict__ order, const uint2 *__restrict__ tbox, const uint64_t *__restrict__ tmask,
              const uint32_t *__restrict__ slot_base, const uint32_t *__restrict__ tile_base,
              const uint32_t *__restrict__ tiles_x, uint32_t n_views, uint32_t tile_passes, uint32_t capacity,
              const FrameCounters *counters, uint32_t *__restrict__ pair_keys, uint32_t *__restrict__ pair_vals,
              const uint64_t *__restrict__ chunk_first, uint32_t *__restrict__ tile_hist) {
  __shared__ uint32_t s_start[PE_CHUNK + 1];
  __shared__ uint16_t s_own[PE_WIN];                 // a window's owners: local rank + 1
  __shared__ uint32_t s_wmax[PE_THREADS / SL_WAVE];
  __shared__ uint2 s_box[PE_CHUNK];
  __shared__ uint32_t s_g[PE_CHUNK];
  __shared__ uint64_t s_mask[MASKED ? PE_CHUNK : 1];
  constexpr uint32_t DIG = 1u << BITS;
  __shared__ uint32_t s_hist[3 * DIG];
  __shared__ uint32_t s_wsum[PE_THREADS / SL_WAVE];

  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t V = __hip_atomic_load(&counters->visible, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t c = blockIdx.x, r0 = c * PE_CHUNK;
  if (r0 >= V) return;
  const uint32_t tb0 = tile_base[0], tx0v = tiles_x[0];   // the only view of a single frame
  for (uint32_t k = tid; k < 3u * DIG; k += PE_THREADS) s_hist[k] = 0u;
  uint32_t cnt[PE_RPT], first[PE_RPT], sum = 0;
#pragma unroll
  for (int k = 0; k < PE_RPT; k++) {
    const uint32_t lr = tid * PE_RPT + k;
    uint32_t g;
    uint2 b;
    uint64_t m;
    cnt[k] = pair_count<MASKED>(order, tbox, tmask, r0 + lr, V, g, b, m);
    if (MASKED) s_mask[lr] = m;
    s_box[lr] = b;
    s_g[lr] = g;
    sum += cnt[k];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(sum, s_wsum, &total);
#pragma unroll
  for (int k = 0; k < PE_RPT; k++) { first[k] = run; s_start[tid * PE_RPT + k] = run; run += cnt[k]; }
  if (tid == 0) s_start[PE_CHUNK] = total;
  const uint64_t base = chunk_first[c];
  uint32_t carry = 0;                                   // the owner (local rank + 1) of the pair before the window
  for (uint32_t w0 = 0; w0 < total; w0 += PE_WIN) {
    for (uint32_t p = tid; p < PE_WIN; p += PE_THREADS) s_own[p] = 0;
    __syncthreads();                                    // (also: the chunk's ranks are in LDS)
#pragma unroll
    for (int k = 0; k < PE_RPT; k++)                   // a rank marks the position of its first pair
      if (cnt[k] != 0u && first[k] >= w0 && first[k] - w0 < PE_WIN)
        s_own[first[k] - w0] = (uint16_t)(tid * PE_RPT + k + 1u);
    __syncthreads();
    // inclusive max-scan of the marks (ranks ascend with their first pair, so the running maximum is the owner):
    // each thread scans PER consecutive entries, then the threads' maxima are scanned across the workgroup
    constexpr uint32_t PER = PE_WIN / PE_THREADS;
    uint32_t own[PER], mx = 0;
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
      mx = max(mx, (uint32_t)s_own[PER * tid + i]);
      own[i] = mx;
    }
    uint32_t xm = mx;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(xm, o);
      if (lane >= (uint32_t)o) xm = max(xm, y);
    }
    if (lane == 63u) s_wmax[w] = xm;
    const uint32_t prev_lane = __shfl_up(xm, 1);
    __syncthreads();
    uint32_t pre = max(carry, lane > 0u ? prev_lane : 0u);
#pragma unroll
    for (int k = 0; k < PE_THREADS / SL_WAVE; k++) pre = (uint32_t)k < w ? max(pre, s_wmax[k]) : pre;
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) s_own[PER * tid + i] = (uint16_t)max(pre, own[i]);
    __syncthreads();
    carry = s_own[PE_WIN - 1];
    const uint32_t w1 = min(total, w0 + PE_WIN);
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
      const uint32_t j = w0 + tid + PE_THREADS * i;
      if (j >= w1) break;
      const uint32_t lo = (uint32_t)s_own[j - w0] - 1u;
      const uint2 b = s_box[lo];
      const uint32_t local = MASKED ? nth_set_bit(s_mask[lo], j - s_start[lo]) : j - s_start[lo];
      const uint32_t tx0 = b.x & 0xffffu, ty0 = b.x >> 16;
      const uint32_t bw = (b.y & 0xffffu) - tx0 + 1u;
      // local / bw and local % bw through the float reciprocal (local < 2^24: the quotient is off by at most one)
      uint32_t qd = (uint32_t)((float)local * __builtin_amdgcn_rcpf((float)bw));
      int32_t rm = (int32_t)local - (int32_t)(qd * bw);
      if (rm < 0) { qd--; rm += (int32_t)bw; } else if (rm >= (int32_t)bw) { qd++; rm -= (int32_t)bw; }
      const uint32_t ty = ty0 + qd, tx = tx0 + (uint32_t)rm;
      uint32_t tile;
      if (n_views > 1u) {
        const uint32_t vw = view_of(slot_base, n_views, s_g[lo]);
        tile = tile_base[vw] + ty * tiles_x[vw] + tx;
      } else {
        tile = tb0 + ty * tx0v + tx;
      }
      const uint64_t dst = base + j;
      const bool keep = dst < capacity;
      if (keep) {
        pair_keys[dst] = tile;
        pair_vals[dst] = s_g[lo];
      }
      // tile-id digit counts for the tile sort: a digit the wave's kept pairs all share (the high digits, nearly
      // always: a wave's pairs lie in a few neighbouring tile rows) is one LDS add of the count; otherwise each lane
      // adds its own (low digits of neighbouring pairs differ, so the adds rarely meet on one address)
      const uint64_t kept = __ballot(keep);
      const bool leader = keep && (kept & lanemask_lt()) == 0ull;
#pragma unroll
      for (uint32_t pass = 0; pass < 3u; pass++) {
        if (pass >= tile_passes) break;
        const uint32_t d = (tile >> (BITS * pass)) & (DIG - 1u);
        const uint32_t d_lead = __builtin_amdgcn_readlane(d, kept ? (uint32_t)__builtin_ctzll(kept) : 0u);
        if (__ballot(keep && d != d_lead) == 0ull) {
          if (leader) atomicAdd(&s_hist[pass * DIG + d_lead], (uint32_t)__popcll(kept));
        } else if (keep) {
          atomicAdd(&s_hist[pass * DIG + d], 1u);
        }
      }
    }
    __syncthreads();                                    // the window's owners are read before the next clear
  }
  __syncthreads();
  for (uint32_t k = tid; k < DIG * tile_passes; k += PE_THREADS)
    if (s_hist[k]) atomicAdd(&tile_hist[k], s_hist[k]);
}

// Pairs per view of a batch: the sum of its tiles' range lengths (one workgroup per view).
extern "C" __global__ void __launch_bounds__(256)
sl_view_pairs(const uint2 *__restrict__ ranges, const uint32_t *__restrict__ tile_base,
              ViewCounters *__restrict__ view_counters) {
  __shared__ uint32_t s_tmp[256 / SL_WAVE];
  const uint32_t v = blockIdx.x;
  uint32_t n = 0;
  for (uint32_t t = tile_base[v] + threadIdx.x; t < tile_base[v + 1]; t += 256u) n += ranges[t].y - ranges[t].x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
  if (lane_id() == 0) s_tmp[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) view_counters[v].pairs = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
}

// Per-tile [start, end) in the tile-sorted pair list; ranges are zero on entry (empty tiles stay [0, 0)).
extern "C" __global__ void __launch_bounds__(256)
sl_tile_ranges(const uint32_t *__restrict__ keys, const FrameCounters *counters, uint2 *__restrict__ ranges) {
  const uint32_t n = __hip_atomic_load(&counters->pairs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint32_t t = keys[i];
    if (i == 0u || keys[i - 1u] != t) ranges[t].x = i;
    if (i == n - 1u || keys[i + 1u] != t) ranges[t].y = i + 1u;
  }
}

// Composite launch order: tiles by pair count, heaviest first, in half-octave buckets (one workgroup). The
// hardware hands out workgroups in launch order as earlier ones retire, so heavy tiles start first and the
// light ones fill in behind them (greedy longest-first), and consecutive workgroups go to different XCDs, so
// each XCD gets an even share of every cost class. Order within a bucket is arbitrary: a tile's pixels do not
// depend on when it runs.
extern "C" __global__ void __launch_bounds__(256)
sl_tile_order(const uint2 *__restrict__ ranges, uint32_t ntiles, uint32_t *__restrict__ order) {
  __shared__ uint32_t s_cnt[64];
  __shared__ uint32_t s_next[64];
  if (threadIdx.x < 64u) s_cnt[threadIdx.x] = 0u;
  __syncthreads();
  auto bucket = [](uint32_t c) -> uint32_t {
    if (c == 0u) return 63u;
    const uint32_t msb = 31u - (uint32_t)__clz(c);
    const uint32_t half = msb > 0u ? (c >> (msb - 1u)) & 1u : 0u;
    return 62u - min(61u, 2u * msb + half);         // heavier -> smaller bucket index
  };
  for (uint32_t t = threadIdx.x; t < ntiles; t += 256u) {
    const uint2 r = ranges[t];
    atomicAdd(&s_cnt[bucket(r.y - r.x)], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int b = 0; b < 64; b++) { s_next[b] = run; run += s_cnt[b]; }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < ntiles; t += 256u) {
    const uint2 r = ranges[t];
    order[atomicAdd(&s_next[bucket(r.y - r.x)], 1u)] = t;
  }
}
