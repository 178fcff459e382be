/*
 * k_prims.hip -- device-wide primitives written for wave64 / gfx950:
 *   - exclusive scan (block tiles of 256 threads x 8 items, wave64 shuffle
 *     scans, recursive over tile sums), replacing the reference's Thrust
 *     inclusive_scan (AlphaComplexParallel.cu:1207, :1287, :1389) and the
 *     lower/upper_bound CSR construction (:1179-1193, :1260-1263);
 *   - stable LSD radix sort of (key, value) pairs with 8-bit digits, stable
 *     in-tile ranking through 64-bit ballots, replacing
 *     thrust::sort_by_key of the cell keys (AlphaComplexParallel.cu:1157).
 */
#include "pac_internal.h"

namespace pac {

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

/* block-wide exclusive scan of one value per thread; returns block total */
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v,
                                                        uint64_t* total) {
  __shared__ uint64_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t inc = wave_incl_scan_u64(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t base = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < kBlock / 64; ++q) {
    if (q < w) base += wsum[q];
    tot += wsum[q];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

template <class TIn, bool kFlag>
__device__ __forceinline__ uint64_t load_val(const TIn* in, int64_t i,
                                             int64_t n) {
  if (i >= n) return 0;
  if (kFlag) return in[i] != 0 ? 1 : 0;
  return (uint64_t)in[i];
}

/* Tiles of kTile items are read and written striped across the block
 * (coalesced), and transposed through LDS so that each thread scans
 * kScanItems consecutive items. */
template <class TIn, bool kFlag>
__global__ __launch_bounds__(kBlock) void k_tile_sums(const TIn* in, Cnt cn,
                                                      uint64_t* sums) {
  const int64_t n = cn.get();
  int64_t base = (int64_t)blockIdx.x * kTile;
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q)
    s += load_val<TIn, kFlag>(in, base + (int64_t)q * kBlock + threadIdx.x, n);
  uint64_t tot;
  block_excl_scan_u64(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class TIn, bool kFlag, class TOut>
__global__ __launch_bounds__(kBlock) void k_tile_scan(const TIn* in, Cnt cn,
                                                      const uint64_t* tile_off,
                                                      TOut* out,
                                                      uint64_t* total64) {
  __shared__ uint64_t tile[kTile];
  const int64_t n = cn.get();
  int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int t = q * kBlock + threadIdx.x;
    tile[t] = load_val<TIn, kFlag>(in, base + t, n);
  }
  __syncthreads();
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = tile[threadIdx.x * kScanItems + q];
    s += v[q];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan_u64(s, &tot) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    tile[threadIdx.x * kScanItems + q] = ex;
    ex += v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int t = q * kBlock + threadIdx.x;
    if (base + t < n) out[base + t] = (TOut)tile[t];
  }
  /* the tile holding position n (the last tile when n ends a tile) writes
   * the grand total at out[n] */
  const bool holds = (n >= base && n < base + kTile) ||
                     (blockIdx.x == gridDim.x - 1 && n == base + kTile);
  if (holds && threadIdx.x == kBlock - 1) {
    out[n] = (TOut)ex;
    if (total64) *total64 = ex;
  }
}

size_t scan_tmp_bytes(int64_t n) {
  size_t b = 0;
  int64_t m = n;
  while (m > kTile) {
    int64_t t = (m + kTile - 1) / kTile;
    b += (size_t)(t + 1) * 2 * sizeof(uint64_t);
    m = t;
  }
  return b + 256;
}

template <class TOut>
__global__ void k_store(TOut* out, TOut v) { *out = v; }

/* Exclusive scan of up to kSumsMax tile sums by one block of 1024
 * threads (out[m] = total): each thread sums a run of consecutive values,
 * the block scans the 1024 run sums (wave64 shuffles, then the 16 wave
 * totals), and each thread writes its run's prefixes. */
constexpr int kSumsBlock = 1024;
constexpr int64_t kSumsMax = (int64_t)kSumsBlock * 64;
__global__ __launch_bounds__(kSumsBlock) void k_scan_sums(const uint64_t* in,
                                                          int64_t m,
                                                          uint64_t* out,
                                                          uint64_t* total64) {
  __shared__ uint64_t wsum[kSumsBlock / 64];
  const int per = (int)((m + kSumsBlock - 1) / kSumsBlock); /* <= 64 */
  const int64_t b = (int64_t)threadIdx.x * per;
  /* the run in batches of 16 values held in registers: a batch's loads
   * are all issued before any is used (one dependent round trip per
   * batch, not per value) */
  constexpr int kB = 16;
  uint64_t run = 0;
  for (int k0 = 0; k0 < per; k0 += kB) {
    uint64_t v[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k)
      v[k] = (k0 + k < per && b + k0 + k < m) ? in[b + k0 + k] : 0;
#pragma unroll
    for (int k = 0; k < kB; ++k) run += v[k];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_scan_u64(run);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t ex = inc - run;
  for (int q = 0; q < w; ++q) ex += wsum[q];
  for (int k0 = 0; k0 < per; k0 += kB) {
    uint64_t v[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k)
      v[k] = (k0 + k < per && b + k0 + k < m) ? in[b + k0 + k] : 0;
#pragma unroll
    for (int k = 0; k < kB; ++k)
      if (k0 + k < per && b + k0 + k < m) {
        out[b + k0 + k] = ex;
        ex += v[k];
      }
  }
  if (threadIdx.x == kSumsBlock - 1) {
    out[m] = ex;
    if (total64) *total64 = ex;
  }
}

/* recursive scan over uint64 tile sums */
static void scan_u64_rec(const uint64_t* in, uint64_t* out, int64_t n,
                         uint64_t* tmp, hipStream_t s);

template <class TIn, bool kFlag, class TOut = uint64_t>
static void scan_impl(const TIn* in, TOut* out, Cnt n, uint64_t* tmp,
                      hipStream_t s, uint64_t* total64 = nullptr) {
  if (n.cap <= 0) {
    hipLaunchKernelGGL(k_store<TOut>, dim3(1), dim3(1), 0, s, out, (TOut)0);
    if (total64)
      hipLaunchKernelGGL(k_store<uint64_t>, dim3(1), dim3(1), 0, s, total64,
                         (uint64_t)0);
    return;
  }
  int64_t tiles = (n.cap + kTile - 1) / kTile;
  if (tiles == 1) {
    hipLaunchKernelGGL((k_tile_scan<TIn, kFlag, TOut>), dim3(1), dim3(kBlock),
                       0, s, in, n, (const uint64_t*)nullptr, out, total64);
    return;
  }
  uint64_t* sums = tmp;
  uint64_t* offs = tmp + (tiles + 1);
  hipLaunchKernelGGL((k_tile_sums<TIn, kFlag>), dim3((unsigned)tiles),
                     dim3(kBlock), 0, s, in, n, sums);
  if (tiles <= kSumsMax) /* the tile sums in one block: three launches */
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumsBlock), 0, s,
                       (const uint64_t*)sums, tiles, offs, (uint64_t*)nullptr);
  else
    scan_u64_rec(sums, offs, tiles, tmp + 2 * (tiles + 1), s);
  hipLaunchKernelGGL((k_tile_scan<TIn, kFlag, TOut>), dim3((unsigned)tiles),
                     dim3(kBlock), 0, s, in, n, (const uint64_t*)offs, out,
                     total64);
}

static void scan_u64_rec(const uint64_t* in, uint64_t* out, int64_t n,
                         uint64_t* tmp, hipStream_t s) {
  scan_impl<uint64_t, false>(in, out, n, tmp, s);
}

void scan_u32_to_u64(const uint32_t* in, uint64_t* out, Cnt n, void* tmp,
                     size_t, hipStream_t s, uint64_t* total64) {
  scan_impl<uint32_t, false>(in, out, n, (uint64_t*)tmp, s, total64);
}
void scan_u64_to_u64(const uint64_t* in, uint64_t* out, Cnt n, void* tmp,
                     size_t, hipStream_t s, uint64_t* total64) {
  scan_impl<uint64_t, false>(in, out, n, (uint64_t*)tmp, s, total64);
}
/* Flag scans (u8 flags, nonzero = 1 -> u32 positions), the scan every
 * compaction runs: a thread owns kScanItems (8) consecutive flags, read as
 * one 8-byte word, and writes their 8 positions as two 16-byte stores;
 * 32-bit arithmetic, no LDS transpose.  Same tiles, tile sums and totals
 * as the generic scan (flag arrays and positions are pool buffers, aligned
 * far beyond 16 bytes, and a tile starts at a multiple of kTile items). */
static_assert(kScanItems == 8, "a thread reads its flags as one 8-byte word");
__device__ __forceinline__ unsigned nz_bytes(uint32_t x) {
  /* high bit of each byte set iff the byte is nonzero */
  const uint32_t y = ((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x;
  return (unsigned)__popc(y & 0x80808080u);
}
__device__ __forceinline__ uint2 load_flags8(const uint8_t* in, int64_t i0,
                                             int64_t n) {
  if (i0 + 8 <= n) return *reinterpret_cast<const uint2*>(in + i0);
  uint32_t w[2] = {0u, 0u};
  for (int q = 0; q < 8; ++q)
    if (i0 + q < n) w[q >> 2] |= (uint32_t)in[i0 + q] << (8 * (q & 3));
  return make_uint2(w[0], w[1]);
}
__global__ __launch_bounds__(kBlock) void k_flag_sums(const uint8_t* in,
                                                      Cnt cn, uint64_t* sums) {
  const int64_t n = cn.get();
  const int64_t i0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * 8;
  const uint2 f = load_flags8(in, i0, n);
  uint64_t tot;
  block_excl_scan_u64(nz_bytes(f.x) + nz_bytes(f.y), &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ __launch_bounds__(kBlock) void k_flag_scan(const uint8_t* in, Cnt cn,
                                                      const uint64_t* tile_off,
                                                      uint32_t* out,
                                                      uint64_t* total64) {
  const int64_t n = cn.get();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const int64_t i0 = base + (int64_t)threadIdx.x * 8;
  const uint2 f = load_flags8(in, i0, n);
  uint64_t tot;
  const uint64_t ex = block_excl_scan_u64(nz_bytes(f.x) + nz_bytes(f.y), &tot) +
                      (tile_off ? tile_off[blockIdx.x] : 0);
  uint32_t p[8];
  uint32_t e = (uint32_t)ex;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    p[q] = e;
    e += ((q < 4 ? f.x : f.y) >> (8 * (q & 3)) & 0xffu) != 0;
  }
  if (i0 + 8 <= n) {
    uint4* o = reinterpret_cast<uint4*>(out + i0);
    o[0] = make_uint4(p[0], p[1], p[2], p[3]);
    o[1] = make_uint4(p[4], p[5], p[6], p[7]);
  } else {
    for (int q = 0; q < 8; ++q)
      if (i0 + q < n) out[i0 + q] = p[q];
  }
  /* the tile holding position n (the last tile when n ends a tile) writes
   * the grand total at out[n] */
  const bool holds = (n >= base && n < base + kTile) ||
                     (blockIdx.x == gridDim.x - 1 && n == base + kTile);
  if (holds && threadIdx.x == 0) {
    const uint64_t t = tot + (tile_off ? tile_off[blockIdx.x] : 0);
    out[n] = (uint32_t)t;
    if (total64) *total64 = t;
  }
}
void scan_u8flag_to_u32(const uint8_t* in, uint32_t* out, Cnt n, void* tmp,
                        size_t, hipStream_t s, uint64_t* total64) {
  if (n.cap <= 0) {
    scan_impl<uint8_t, true, uint32_t>(in, out, n, (uint64_t*)tmp, s, total64);
    return;
  }
  const int64_t tiles = (n.cap + kTile - 1) / kTile;
  uint64_t* t = (uint64_t*)tmp;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(kBlock), 0, s, in, n,
                       (const uint64_t*)nullptr, out, total64);
    return;
  }
  uint64_t* sums = t;
  uint64_t* offs = t + (tiles + 1);
  hipLaunchKernelGGL(k_flag_sums, dim3((unsigned)tiles), dim3(kBlock), 0, s, in,
                     n, sums);
  if (tiles <= kSumsMax)
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumsBlock), 0, s,
                       (const uint64_t*)sums, tiles, offs, (uint64_t*)nullptr);
  else
    scan_u64_rec(sums, offs, tiles, t + 2 * (tiles + 1), s);
  hipLaunchKernelGGL(k_flag_scan, dim3((unsigned)tiles), dim3(kBlock), 0, s, in,
                     n, (const uint64_t*)offs, out, total64);
}

/* Compaction by byte flags with no position per entry: the ladder's tail
 * computes each tile's offset (flag_tile_offsets: k_flag_sums, then the
 * scan of the tile sums, whose total is the count), and the compaction
 * kernels rank each tile's flags themselves with the block scan of
 * k_flag_scan and write only the kept entries.  Against a position array
 * (k_flag_scan, then one thread per flag reading flag and position) this
 * writes and reads no 4 bytes per entry and launches 8x fewer waves. */
__global__ void k_copy_u64(const uint64_t* src, uint64_t* dst) { *dst = *src; }
const uint64_t* flag_tile_offsets(const uint8_t* in, Cnt n, void* tmp,
                                  hipStream_t s, uint64_t* total64) {
  uint64_t* t = (uint64_t*)tmp;
  const int64_t tiles = (n.cap + kTile - 1) / kTile;
  if (tiles <= 0) {
    if (total64)
      hipLaunchKernelGGL(k_store<uint64_t>, dim3(1), dim3(1), 0, s, total64,
                         (uint64_t)0);
    return t;
  }
  uint64_t* sums = t;
  uint64_t* offs = t + (tiles + 1);
  hipLaunchKernelGGL(k_flag_sums, dim3((unsigned)tiles), dim3(kBlock), 0, s, in,
                     n, sums);
  if (tiles <= kSumsMax) {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumsBlock), 0, s,
                       (const uint64_t*)sums, tiles, offs, total64);
  } else {
    scan_u64_rec(sums, offs, tiles, t + 2 * (tiles + 1), s);
    if (total64)
      hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(1), 0, s,
                         (const uint64_t*)(offs + tiles), total64);
  }
  return offs;
}
/* the kept entries of this block's tile, in order: emit(i, p), p = the
 * entry's rank among all kept entries */
template <class F>
__device__ __forceinline__ void tile_kept(const uint8_t* flag, int64_t n,
                                          const uint64_t* offs, F&& emit) {
  const int64_t i0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * 8;
  const uint2 f = load_flags8(flag, i0, n);
  uint64_t tot;
  uint64_t p = block_excl_scan_u64(nz_bytes(f.x) + nz_bytes(f.y), &tot) +
               offs[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if ((((q < 4 ? f.x : f.y) >> (8 * (q & 3))) & 0xffu) != 0) emit(i0 + q, p++);
}
__global__ __launch_bounds__(kBlock) void k_tile_compact_u4(
    Cnt cn, const uint4* in, const uint8_t* flag, const uint64_t* offs,
    uint4* out, uint64_t ocap) {
  tile_kept(flag, cn.get(), offs, [&](int64_t i, uint64_t p) {
    if (p < ocap) out[p] = in[i];
  });
}
__global__ __launch_bounds__(kBlock) void k_tile_compact_ids(
    Cnt cn, const uint8_t* flag, const uint64_t* offs, uint32_t* out,
    uint64_t ocap) {
  tile_kept(flag, cn.get(), offs, [&](int64_t i, uint64_t p) {
    if (p < ocap) out[p] = (uint32_t)i;
  });
}
__global__ __launch_bounds__(kBlock) void k_tile_compact_u8(
    Cnt cn, const uint8_t* flag, const uint64_t* offs, const uint8_t* val,
    uint8_t* out, uint64_t ocap) {
  tile_kept(flag, cn.get(), offs, [&](int64_t i, uint64_t p) {
    if (p < ocap) out[p] = val[i] != 0;
  });
}
__global__ __launch_bounds__(kBlock) void k_tile_emit_edges(
    Cnt cn, const uint32_t* ecol, const uint32_t* eleft, const uint8_t* flag,
    const uint64_t* offs, const uint32_t* perm, int2* out, uint64_t ocap) {
  tile_kept(flag, cn.get(), offs, [&](int64_t i, uint64_t p) {
    if (p < ocap) out[p] = make_int2((int)perm[eleft[i]], (int)perm[ecol[i]]);
  });
}
static dim3 tiles_of(Cnt n) { return dim3((unsigned)((n.cap + kTile - 1) / kTile)); }
void launch_compact_tets(Cnt n, const uint4* in, const uint8_t* flag,
                         const uint64_t* offs, uint4* out, PipeLaunch L,
                         uint64_t ocap) {
  if (n.cap <= 0) return;
  hipLaunchKernelGGL(k_tile_compact_u4, tiles_of(n), dim3(kBlock), 0, L.s, n,
                     in, flag, offs, out, ocap);
}
void launch_compact_ids(Cnt n, const uint8_t* flag, const uint64_t* offs,
                        uint32_t* out, PipeLaunch L, uint64_t ocap) {
  if (n.cap <= 0) return;
  hipLaunchKernelGGL(k_tile_compact_ids, tiles_of(n), dim3(kBlock), 0, L.s, n,
                     flag, offs, out, ocap);
}
void launch_compact_u8(Cnt n, const uint8_t* flag, const uint64_t* offs,
                       const uint8_t* val, uint8_t* out, PipeLaunch L,
                       uint64_t ocap) {
  if (n.cap <= 0) return;
  hipLaunchKernelGGL(k_tile_compact_u8, tiles_of(n), dim3(kBlock), 0, L.s, n,
                     flag, offs, val, out, ocap);
}
void launch_emit_edges(Cnt ne, const uint32_t* ecol, const uint32_t* eleft,
                       const uint8_t* flag, const uint64_t* offs,
                       const uint32_t* perm, int32_t* out, PipeLaunch L,
                       uint64_t ocap) {
  if (ne.cap <= 0) return;
  hipLaunchKernelGGL(k_tile_emit_edges, tiles_of(ne), dim3(kBlock), 0, L.s, ne,
                     ecol, eleft, flag, offs, perm, (int2*)out, ocap);
}

/* ------------------------------------------------------------------ */
/* radix sort                                                          */
/* ------------------------------------------------------------------ */
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* keys,
                                                       int64_t n, int shift,
                                                       int64_t ntiles,
                                                       uint32_t* hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int64_t i = base + (int64_t)q * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

/* Stable scatter: items are ranked round by round (256 consecutive items
 * per round); inside a wave, lanes with equal digits are found with eight
 * 64-bit ballots; across the four waves and earlier rounds, LDS counters
 * give the offset. */
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t* keys, const uint32_t* vals, int64_t n, int shift,
    int64_t ntiles, const uint64_t* hoff, uint32_t* keys_out,
    uint32_t* vals_out) {
  __shared__ uint32_t running[256];
  __shared__ uint32_t wcnt[kBlock / 64][256];
  __shared__ uint64_t base_of[256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  running[threadIdx.x] = 0;
  base_of[threadIdx.x] = hoff[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  for (int q = 0; q < kBlock / 64; ++q) wcnt[q][threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kTile;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kScanItems; ++r) {
    int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
    bool valid = i < n;
    uint32_t key = valid ? keys[i] : 0u;
    uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    bool leader = valid && (peers & lt_mask) == 0;
    if (leader) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t off = running[d] + rank;
      for (int q = 0; q < w; ++q) off += wcnt[q][d];
      uint64_t dst = base_of[d] + off;
      keys_out[dst] = key;
      vals_out[dst] = vals[i];
    }
    __syncthreads();
    {
      uint32_t add = 0;
      for (int q = 0; q < kBlock / 64; ++q) {
        add += wcnt[q][threadIdx.x];
        wcnt[q][threadIdx.x] = 0;
      }
      running[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

size_t radix_tmp_bytes(int64_t n) {
  int64_t ntiles = (n + kTile - 1) / kTile;
  size_t h = (size_t)256 * ntiles;
  return h * sizeof(uint32_t) + (h + 1) * sizeof(uint64_t) + scan_tmp_bytes(h) + 512;
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                      uint32_t* vals_alt, int64_t n, int bits, void* tmp,
                      size_t, hipStream_t s, bool* result_in_alt) {
  *result_in_alt = false;
  if (n <= 1 || bits <= 0) return;
  int64_t ntiles = (n + kTile - 1) / kTile;
  size_t h = (size_t)256 * ntiles;
  uint32_t* hist = (uint32_t*)tmp;
  uint64_t* hoff = (uint64_t*)(((uintptr_t)(hist + h) + 15) & ~(uintptr_t)15);
  uint64_t* stmp = hoff + h + 1;
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  bool alt = false;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kBlock), 0,
                       s, ki, n, shift, ntiles, hist);
    scan_impl<uint32_t, false>(hist, hoff, (int64_t)h, stmp, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)ntiles), dim3(kBlock),
                       0, s, ki, vi, n, shift, ntiles, hoff, ko, vo);
    uint32_t* t = ki;
    ki = ko;
    ko = t;
    t = vi;
    vi = vo;
    vo = t;
    alt = !alt;
  }
  *result_in_alt = alt;
}

/* ------------------------------------------------------------------ */
/* recorded (fixed-capacity) pipelines                                 */
/* ------------------------------------------------------------------ */
/* Consumer view of QSLOTS sub-queues from their counters; a counter past
 * the capacity (entries were dropped) sets `bit` in *status and the
 * largest counter goes to *fill (fresh: both overwritten and fill[1]
 * cleared, for the ladder's second view; else merged).
 * The counters it read, and *also (if set), are cleared for the next
 * producer, so no separate zeroing launch precedes a ladder. */
__global__ void k_make_view(const QEntry* q, uint32_t cap, unsigned* cnt,
                            QView* v, unsigned* status, unsigned bit,
                            unsigned* fill, unsigned* also, int fresh) {
  if (threadIdx.x != 0) return;
  unsigned c[QSLOTS];
  unsigned mx = 0;
  for (int k = 0; k < QSLOTS; ++k) {
    c[k] = cnt[k];
    cnt[k] = 0;
    mx = c[k] > mx ? c[k] : mx;
  }
  /* an eager ladder's first view (fresh) is empty when its queue
   * overflowed: the ladder runs again with a larger queue, and its
   * consumers need not work through the partial one */
  const bool drop = fresh && mx > cap;
  uint32_t acc = 0;
  for (int k = 0; k < QSLOTS; ++k) {
    v->off[k] = acc;
    acc += drop ? 0u : (c[k] < cap ? c[k] : cap);
  }
  v->off[QSLOTS] = acc;
  v->q = q;
  v->cap = cap;
  if (also)
    for (int k = 0; k < LSLOTS; ++k) also[k] = 0;
  const unsigned b = mx > cap ? bit : 0u;
  if (fresh) { /* first view of a ladder: also resets the second's fill */
    fill[0] = mx;
    fill[1] = 0;
    *status = b;
  } else {
    atomicMax(fill, mx);
    if (b) atomicOr(status, b);
  }
}
void launch_make_view(const QEntry* q, uint32_t cap, unsigned* cnt, QView* v,
                      unsigned* status, unsigned bit, unsigned* fill,
                      unsigned* also, bool fresh, hipStream_t s) {
  hipLaunchKernelGGL(k_make_view, dim3(1), dim3(64), 0, s, q, cap, cnt, v,
                     status, bit, fill, also, (int)fresh);
}

/* An eager ladder's state in one buffer for one read-back: the
 * sub-queue counts of its last queue (then cleared for the next producer)
 * and its count, beside the status and fill its first view left */
__global__ void k_ladder_state(unsigned* cnt, const uint64_t* count,
                               unsigned* lad) {
  const int t = threadIdx.x;
  if (t < QSLOTS) {
    lad[kLadCnt + t] = cnt[t];
    cnt[t] = 0;
  }
  if (t == 0) *reinterpret_cast<uint64_t*>(lad + kLadCount) = *count;
}
void launch_ladder_state(unsigned* cnt, const uint64_t* count, unsigned* lad,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_ladder_state, dim3(1), dim3(64), 0, s, cnt, count, lad);
}

__global__ void k_cap_check(const uint64_t* v, uint64_t cap, unsigned* status,
                            unsigned bit) {
  if (threadIdx.x == 0 && *v > cap) atomicOr(status, bit);
}
void launch_cap_check(const uint64_t* v, uint64_t cap, unsigned* status,
                      unsigned bit, hipStream_t s) {
  hipLaunchKernelGGL(k_cap_check, dim3(1), dim3(64), 0, s, v, cap, status, bit);
}

/* bounds of the current input against those the pipeline was planned for:
 * any difference in the fields the plan depends on (corners, radii,
 * largest |coordinate|), a non-finite input or a finer input lsb than
 * the exact scale covers sets `bit` */
__device__ __forceinline__ bool same_f64(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
__global__ void k_plan_check(const Bounds* now, const Bounds* ref, int64_t m,
                             int scale, unsigned* status, unsigned bit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Bounds a = now[i], b = ref[i];
  const bool ok = same_f64(a.minx, b.minx) && same_f64(a.miny, b.miny) &&
                  same_f64(a.minz, b.minz) && same_f64(a.maxx, b.maxx) &&
                  same_f64(a.maxy, b.maxy) && same_f64(a.maxz, b.maxz) &&
                  same_f64(a.max_raw_r, b.max_raw_r) &&
                  same_f64(a.max_r, b.max_r) && same_f64(a.max_abs, b.max_abs) &&
                  a.nonfinite == 0 && a.shift <= scale;
  if (!ok) atomicOr(status, bit);
}
void launch_plan_check(const Bounds* now, const Bounds* ref, int64_t m,
                       int scale, unsigned* status, unsigned bit,
                       hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_plan_check, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, now, ref, m, scale, status, bit);
}

}  // namespace pac
