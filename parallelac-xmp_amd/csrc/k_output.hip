/*
 * k_output.hip -- per-simplex alpha values and canonical row order.
 *
 * Alpha values: rho (squared orthoradius minus nothing: the power of the
 * smallest orthosphere, the quantity the reference compares with alpha and
 * stores in float4 .w, AlphaComplexParallel.cu:370, :603) of every output
 * simplex, evaluated in FP64 from the FP32 balls with the closed forms the
 * filters use:
 *   edge      (A^2)/(4|u|^2) - w_i,                A = |u|^2 + w_i - w_j
 *   triangle  (A^2|v|^2 - 2AB u.v + B^2|u|^2)/(4|n|^2) - w_i
 *   tet       |N|^2/(4 D^2) - w_i
 * Canonical order: vertices of each row sorted by original id, rows sorted
 * lexicographically (LSD radix over the columns).
 */
#include "pac_dev.h"

namespace pac {

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_rho(int64_t m, const int32_t* rows,
                                                const float4* balls,
                                                double* rho) {
  int64_t id = gtid();
  if (id >= m) return;
  const int32_t* r = rows + (int64_t)(DIM + 1) * id;
  float4 a = balls[r[0]], b = balls[r[1]];
  double ax = a.x, ay = a.y, az = a.z, wa = (double)a.w * a.w;
  double ux = b.x - ax, uy = b.y - ay, uz = b.z - az;
  double uu = ux * ux + uy * uy + uz * uz;
  double A = uu + wa - (double)b.w * b.w;
  double out;
  if (DIM == 1) {
    out = A * A / (4.0 * uu) - wa;
  } else if (DIM == 2) {
    float4 c = balls[r[2]];
    double vx = c.x - ax, vy = c.y - ay, vz = c.z - az;
    double vv = vx * vx + vy * vy + vz * vz, uv = ux * vx + uy * vy + uz * vz;
    double B = vv + wa - (double)c.w * c.w;
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    double nn = nx * nx + ny * ny + nz * nz;
    out = (A * A * vv - 2.0 * A * B * uv + B * B * uu) / (4.0 * nn) - wa;
  } else {
    float4 c = balls[r[2]], d = balls[r[3]];
    double vx = c.x - ax, vy = c.y - ay, vz = c.z - az;
    double tx = d.x - ax, ty = d.y - ay, tz = d.z - az;
    double B = vx * vx + vy * vy + vz * vz + wa - (double)c.w * c.w;
    double C = tx * tx + ty * ty + tz * tz + wa - (double)d.w * d.w;
    double vtx = vy * tz - vz * ty, vty = vz * tx - vx * tz, vtz = vx * ty - vy * tx;
    double tux = ty * uz - tz * uy, tuy = tz * ux - tx * uz, tuz = tx * uy - ty * ux;
    double uvx = uy * vz - uz * vy, uvy = uz * vx - ux * vz, uvz = ux * vy - uy * vx;
    double D = ux * vtx + uy * vty + uz * vtz;
    double Nx = A * vtx + B * tux + C * uvx, Ny = A * vty + B * tuy + C * uvy,
           Nz = A * vtz + B * tuz + C * uvz;
    out = (Nx * Nx + Ny * Ny + Nz * Nz) / (4.0 * D * D) - wa;
  }
  rho[id] = out;
}
void launch_rho(int dim, int64_t m, const int32_t* rows, const float4* balls,
                double* rho, hipStream_t s) {
  if (m <= 0) return;
  dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  if (dim == 1)
    hipLaunchKernelGGL(k_rho<1>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
  else if (dim == 2)
    hipLaunchKernelGGL(k_rho<2>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
  else
    hipLaunchKernelGGL(k_rho<3>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
}

/* sort the vertices of each row ascending (rows of k = DIM+1 ids) */
template <int K>
__global__ __launch_bounds__(kBlock) void k_sort_rows(int64_t m, int32_t* rows) {
  int64_t id = gtid();
  if (id >= m) return;
  int32_t v[K];
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = rows[K * id + q];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b + 1 < K - a; ++b)
      if (v[b] > v[b + 1]) {
        int32_t t = v[b];
        v[b] = v[b + 1];
        v[b + 1] = t;
      }
#pragma unroll
  for (int q = 0; q < K; ++q) rows[K * id + q] = v[q];
}
/* key[i] = rows[K*idx[i] + col]; val stays idx */
__global__ __launch_bounds__(kBlock) void k_col_key(int64_t m, int K,
                                                    const int32_t* rows,
                                                    const uint32_t* idx,
                                                    int col, uint32_t* key) {
  int64_t i = gtid();
  if (i >= m) return;
  key[i] = (uint32_t)rows[(int64_t)K * idx[i] + col];
}
__global__ __launch_bounds__(kBlock) void k_iota(int64_t m, uint32_t* idx) {
  int64_t i = gtid();
  if (i < m) idx[i] = (uint32_t)i;
}
__global__ __launch_bounds__(kBlock) void k_gather_rows(int64_t m, int K,
                                                        const int32_t* in,
                                                        const uint32_t* idx,
                                                        int32_t* out) {
  int64_t i = gtid();
  if (i >= m) return;
  for (int q = 0; q < K; ++q) out[(int64_t)K * i + q] = in[(int64_t)K * idx[i] + q];
}

/* canonical order of m rows of K ids (< 2^idbits) in place; tmp buffers are
 * provided by the caller: idx/key/idx2/key2 (m each), radix tmp, rows copy */
void canonical_rows(int K, int64_t m, int32_t* rows, int idbits,
                    uint32_t* idx, uint32_t* key, uint32_t* idx2,
                    uint32_t* key2, void* rtmp, size_t rtb, int32_t* copy,
                    hipStream_t s) {
  if (m <= 0) return;
  dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  if (K == 2)
    hipLaunchKernelGGL(k_sort_rows<2>, gr, dim3(kBlock), 0, s, m, rows);
  else if (K == 3)
    hipLaunchKernelGGL(k_sort_rows<3>, gr, dim3(kBlock), 0, s, m, rows);
  else
    hipLaunchKernelGGL(k_sort_rows<4>, gr, dim3(kBlock), 0, s, m, rows);
  hipLaunchKernelGGL(k_iota, gr, dim3(kBlock), 0, s, m, idx);
  for (int col = K - 1; col >= 0; --col) {
    hipLaunchKernelGGL(k_col_key, gr, dim3(kBlock), 0, s, m, K,
                       (const int32_t*)rows, (const uint32_t*)idx, col, key);
    bool alt = false;
    radix_sort_pairs(key, idx, key2, idx2, m, idbits, rtmp, rtb, s, &alt);
    if (alt) {
      uint32_t* t = idx;
      idx = idx2;
      idx2 = t;
      t = key;
      key = key2;
      key2 = t;
    }
  }
  hipLaunchKernelGGL(k_gather_rows, gr, dim3(kBlock), 0, s, m, K,
                     (const int32_t*)rows, (const uint32_t*)idx, copy);
  (void)hipMemcpyAsync(rows, copy, (size_t)m * K * sizeof(int32_t),
                       hipMemcpyDeviceToDevice, s);
}

/* Rows of a batch's list are grouped by molecule in molecule order and the
 * first id of a row lies in its molecule's id range, so "first id >=
 * moff[m]" is monotone over the rows: out[m] = first row of molecule m. */
__global__ __launch_bounds__(kBlock) void k_mol_rows(const int32_t* rows,
                                                    int64_t m, int K,
                                                    const uint32_t* moff,
                                                    int64_t nmol,
                                                    uint64_t* out) {
  int64_t q = gtid();
  if (q > nmol) return;
  uint32_t key = moff[q];
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((uint32_t)rows[(int64_t)K * mid] < key) lo = mid + 1;
    else hi = mid;
  }
  out[q] = (uint64_t)lo;
}
void launch_mol_rows(const int32_t* rows, int64_t m, int K,
                     const uint32_t* moff, int64_t nmol, uint64_t* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_mol_rows, dim3((unsigned)((nmol + 1 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, rows, m, K, moff, nmol, out);
}

}  // namespace pac
