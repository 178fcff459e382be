/*
 * k_chunk.hip -- the small kernels of the chunked (LD) driver.
 *
 * The reference's LD driver walks the cell-sorted balls in chunks and keeps
 * the simplices whose lowest sorted vertex lies in the chunk
 * (AlphaComplexLargeData.cu:1232-1233, :775-778).  Here a chunk is run as a
 * sub-input (the chunk and a halo, in input order) through the ordinary
 * pipeline; these kernels pick the sub-input, find the rows a chunk owns in
 * each output list and map them back to ids of the whole input.  All are
 * one thread per item and HBM-bound; they do not depend on the ball type
 * except the ball gather, which has an overload per type.
 */
#include "pac_dev.h"

namespace pac {
using namespace PAC_V; /* type-independent device helpers */

/* rank[sidx[p]] = p: the cell-sorted position of every input ball */
__global__ __launch_bounds__(kBlock) void k_rank_of(const uint32_t* sidx,
                                                    int64_t n,
                                                    uint32_t* rank) {
  const int64_t p = gtid();
  if (p >= n) return;
  rank[sidx[p]] = (uint32_t)p;
}
void launch_rank_of(const uint32_t* sidx, int64_t n, uint32_t* rank,
                    hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rank_of, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, sidx, n, rank);
}

/* flag[i] = rank[i] in [lo, hi): input balls of a range of sorted ids */
__global__ __launch_bounds__(kBlock) void k_rank_flag(const uint32_t* rank,
                                                      int64_t n, uint32_t lo,
                                                      uint32_t hi,
                                                      uint8_t* flag) {
  const int64_t i = gtid();
  if (i >= n) return;
  const uint32_t r = rank[i];
  flag[i] = r >= lo && r < hi;
}
void launch_rank_flag(const uint32_t* rank, int64_t n, uint32_t lo,
                      uint32_t hi, uint8_t* flag, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rank_flag, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, rank, n, lo, hi, flag);
}

/* the flagged balls, in input order: out[pos[i]] = raw[i], sid[pos[i]] = i */
template <class B>
__global__ __launch_bounds__(kBlock) void k_select_balls(
    const B* raw, const uint8_t* flag, const uint32_t* pos, int64_t n,
    int64_t m, B* out, uint32_t* sid) {
  const int64_t i = gtid();
  if (i >= n || !flag[i]) return;
  const uint32_t k = pos[i];
  if (k >= m) return;
  out[k] = raw[i];
  sid[k] = (uint32_t)i;
}
template <class B>
static void select_balls(const B* raw, const uint8_t* flag,
                         const uint32_t* pos, int64_t n, int64_t m, B* out,
                         uint32_t* sid, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_select_balls<B>),
                     dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, s, raw, flag, pos, n, m, out, sid);
}
void launch_select_balls(const float4* raw, const uint8_t* flag,
                         const uint32_t* pos, int64_t n, int64_t m,
                         float4* out, uint32_t* sid, hipStream_t s) {
  select_balls(raw, flag, pos, n, m, out, sid, s);
}
void launch_select_balls(const double4* raw, const uint8_t* flag,
                         const uint32_t* pos, int64_t n, int64_t m,
                         double4* out, uint32_t* sid, hipStream_t s) {
  select_balls(raw, flag, pos, n, m, out, sid, s);
}

__global__ void k_set2(uint64_t* out, uint64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = v;
    out[1] = v;
  }
}

/* Rows of a list (width w, sub-input ids, reference order) whose first
 * vertex -- the lowest in sorted order -- has a whole-input sorted
 * position in [lo, hi).  Those positions do not decrease along the list,
 * so the rows form one run [out[0], out[1]): each bound is the single row
 * where the position first reaches lo (hi); with none it stays `rows`. */
__global__ __launch_bounds__(kBlock) void k_row_bounds(
    const int32_t* list, int64_t rows, int w, const uint32_t* sid,
    const uint32_t* rank, uint32_t lo, uint32_t hi, uint64_t* out) {
  const int64_t r = gtid();
  if (r >= rows) return;
  const int64_t g = rank[sid[list[r * w]]];
  const int64_t gp = r > 0 ? (int64_t)rank[sid[list[(r - 1) * w]]] : -1;
  if (gp < (int64_t)lo && g >= (int64_t)lo) out[0] = (uint64_t)r;
  if (gp < (int64_t)hi && g >= (int64_t)hi) out[1] = (uint64_t)r;
}
void launch_row_bounds(const int32_t* list, int64_t rows, int w,
                       const uint32_t* sid, const uint32_t* rank, uint32_t lo,
                       uint32_t hi, uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, s, out, (uint64_t)rows);
  if (rows <= 0) return;
  hipLaunchKernelGGL(k_row_bounds,
                     dim3((unsigned)((rows + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, list, rows, w, sid, rank, lo, hi, out);
}

/* out[j] = sid[in[j]]: sub-input ids to whole-input ids */
__global__ __launch_bounds__(kBlock) void k_map_ids(const int32_t* in,
                                                    int64_t cnt,
                                                    const uint32_t* sid,
                                                    int32_t* out) {
  const int64_t j = gtid();
  if (j >= cnt) return;
  out[j] = (int32_t)sid[in[j]];
}
void launch_map_ids(const int32_t* in, int64_t cnt, const uint32_t* sid,
                    int32_t* out, hipStream_t s) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_map_ids, dim3((unsigned)((cnt + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, in, cnt, sid, out);
}

}  // namespace pac
