// (splat, tile) pair emission and per-tile ranges.
//
// A pair is a visible splat and a 16x16 tile whose pixel-centre range meets the splat's extent box
// (shaders/project.wgsl:314). Pairs are emitted in the depth order produced by the depth sort, so after a
// stable sort by tile id every tile's list is that tile's slice of the reference's global back-to-front order.
//
// sl_pairs_emit is one pass over the visible ranks: persistent workgroups take 1024-rank chunks in ticket
// order, scan the chunk's pair counts, get the chunk's global offset by a decoupled look-back (64-bit status
// words), then emit the chunk's pairs cooperatively: lane j of the workgroup writes pair j, finding its rank by a
// binary search over the chunk's offsets in LDS, so the key and value stores are contiguous. The tile-id digit
// histograms for the tile sort are accumulated on the way.
#include "sl_device.h"

#define PE_THREADS 256
#define PE_RPT 4
#define PE_CHUNK (PE_THREADS * PE_RPT)
#define PE_FLAG_AGG (1ull << 62)
#define PE_FLAG_INC (2ull << 62)
#define PE_FLAG_MASK (3ull << 62)
#define PE_VALUE_MASK ((1ull << 62) - 1ull)

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

// order == nullptr: rank r is slot r. Batches: a rank's slot belongs to the view v with slot_base[v] <= slot < slot_base[v + 1]; its pairs' tile ids
// are global, tile_base[v] + ty * tiles_x[v] + tx. tile_passes (1 to 3) digit histograms of BITS-bit digits (the tile
// sort's digit width) are built.
// MASKED (the opacity-map splats, boxes of at most 64 tiles): only the box's tiles whose bit (row * box width +
// column) is set in tmask[slot] get a pair; otherwise every tile of the box does.
template <bool MASKED, uint32_t BITS>
__global__ void __launch_bounds__(PE_THREADS)
sl_pairs_emit(const uint32_t *__restrict__ order, const uint2 *__restrict__ tbox, const uint64_t *__restrict__ tmask,
              const uint32_t *__restrict__ slot_base, const uint32_t *__restrict__ tile_base,
              const uint32_t *__restrict__ tiles_x, uint32_t n_views, uint32_t tile_passes, uint32_t capacity,
              FrameCounters *counters, uint32_t *__restrict__ pair_keys, uint32_t *__restrict__ pair_vals,
              uint64_t *chunk_status, uint32_t *ticket, uint32_t *__restrict__ tile_hist) {
  __shared__ uint32_t s_start[PE_CHUNK + 1];
  __shared__ uint2 s_box[PE_CHUNK];
  __shared__ uint32_t s_g[PE_CHUNK];
  __shared__ uint64_t s_mask[MASKED ? PE_CHUNK : 1];
  constexpr uint32_t DIG = 1u << BITS;
  __shared__ uint32_t s_hist[3 * DIG];
  __shared__ uint32_t s_wsum[PE_THREADS / SL_WAVE];
  __shared__ uint32_t s_chunk;
  __shared__ uint64_t s_excl;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t tb0 = tile_base[0], tx0v = tiles_x[0];   // the only view of a single frame
  const uint32_t V = __hip_atomic_load(&counters->visible, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t num_chunks = (V + PE_CHUNK - 1u) / PE_CHUNK;
  for (uint32_t k = tid; k < 3u * DIG; k += PE_THREADS) s_hist[k] = 0u;

  for (;;) {
    if (tid == 0) s_chunk = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t c = s_chunk;
    if (c >= num_chunks) break;
    const uint32_t r0 = c * PE_CHUNK;
    // per-thread counts of PE_RPT consecutive ranks
    uint32_t cnt[PE_RPT], sum = 0;
#pragma unroll
    for (int k = 0; k < PE_RPT; k++) {
      const uint32_t lr = tid * PE_RPT + k;
      const uint32_t r = r0 + lr;
      uint32_t n = 0;
      uint2 b = make_uint2(0xffffffffu, 0u);
      uint32_t g = 0;
      if (r < V) {
        g = order ? order[r] : r;
        b = tbox[g];
        if (b.x != 0xffffffffu) {
          if (MASKED) {
            const uint64_t m = tmask[g];
            s_mask[lr] = m;
            n = (uint32_t)__popcll(m);
          } else {
            n = ((b.y & 0xffffu) - (b.x & 0xffffu) + 1u) * ((b.y >> 16) - (b.x >> 16) + 1u);
          }
        }
      }
      s_box[lr] = b;
      s_g[lr] = g;
      cnt[k] = n;
      sum += n;
    }
    // block exclusive scan of the per-thread sums
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) s_wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < PE_THREADS / SL_WAVE; k++) {
      before += (uint32_t)k < w ? s_wsum[k] : 0u;
      total += s_wsum[k];
    }
    uint32_t run = before + x - sum;
#pragma unroll
    for (int k = 0; k < PE_RPT; k++) { s_start[tid * PE_RPT + k] = run; run += cnt[k]; }
    if (w == 0) {
      // decoupled look-back over chunks, 64 predecessors per round: lane l reads chunk c - 1 - l - 64 k; the nearest
      // inclusive prefix ends the walk, and every chunk between must have published its aggregate
      uint64_t *st = chunk_status + c;
      uint64_t excl = 0;
      if (c == 0) {
        if (lane == 0) __hip_atomic_store(st, PE_FLAG_INC | (uint64_t)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        if (lane == 0) __hip_atomic_store(st, PE_FLAG_AGG | (uint64_t)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t top = (int64_t)c - 1;
        for (;;) {
          const int64_t j = top - (int64_t)lane;
          const uint64_t sv = j >= 0 ? __hip_atomic_load(chunk_status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : PE_FLAG_INC;   // before chunk 0: an inclusive zero
          const uint64_t fl = sv & PE_FLAG_MASK;
          const uint64_t inc = __ballot(fl == PE_FLAG_INC), none = __ballot(fl == 0ull);
          const uint32_t first = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;   // the nearest inclusive prefix
          const uint64_t before = first >= 64u ? ~0ull : ((2ull << first) - 1ull);
          if (none & before) {                       // a chunk in between has not published yet
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint64_t v = lane <= first ? (sv & PE_VALUE_MASK) : 0ull;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
          excl += v;
          if (first < 64u) break;
          top -= 64;
        }
        if (lane == 0) __hip_atomic_store(st, PE_FLAG_INC | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (lane == 0) {
        s_start[PE_CHUNK] = total;
        s_excl = excl;
        if (c == num_chunks - 1u) {
          const uint64_t all = excl + total;
          const uint32_t kept = all > capacity ? capacity : (uint32_t)all;
          __hip_atomic_store(&counters->pairs, kept, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&counters->overflow, (uint32_t)(all - kept), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    __syncthreads();
    const uint64_t base = s_excl;
    // cooperative emission: lane j writes pair j of the chunk
    uint32_t lo = 0;                                    // the last rank with start <= j, carried: j only grows
    for (uint32_t j = tid; j < total; j += PE_THREADS) {
      // gallop up from the last owner, then halve (s_start[PE_CHUNK] = total > j ends both): a rank's pairs are
      // consecutive, so a lane's owner moves on by a few ranks where pairs are dense (the heavy chunks)
      uint32_t step = 1u, hi = lo + 1u;
      while (s_start[hi] <= j) {
        lo = hi;
        step <<= 1;
        hi = min(lo + step, (uint32_t)PE_CHUNK);
      }
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_start[mid] <= j) lo = mid; else hi = mid;
      }
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
    __syncthreads();
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
extern "C" __global__ void __launch_bounds__(1024)
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
  for (uint32_t t = threadIdx.x; t < ntiles; t += 1024u) {
    const uint2 r = ranges[t];
    atomicAdd(&s_cnt[bucket(r.y - r.x)], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int b = 0; b < 64; b++) { s_next[b] = run; run += s_cnt[b]; }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < ntiles; t += 1024u) {
    const uint2 r = ranges[t];
    order[atomicAdd(&s_next[bucket(r.y - r.x)], 1u)] = t;
  }
}
