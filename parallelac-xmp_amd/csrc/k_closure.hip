/*
 * k_closure.hip -- closure of the complex and output emission.
 *
 * The reference builds the faces of the final tets and the edges of the
 * final triangles, sorts and de-duplicates them and merges them with the
 * surviving dangling simplices (AlphaComplexParallel.cu:1490-1632,
 * :1646-1788; generateTrisInTets / markTrisForDeletion / generateEdgesInTris
 * / markEdgesForDeletion, AlphaComplexLargeData.cu:737-926).  Every such
 * face is itself an alpha simplex, so here each final tet (triangle) marks
 * its faces (edges) in the already sorted alpha lists by CSR lookup; the
 * final lists are then a stable compaction of the alpha lists and come out
 * in the reference's lexicographic order with no sort or merge.
 */
#include "pac_dev.h"

namespace pac {
using namespace PAC_V; /* type-independent device helpers */

/* out[i] = a[i] | b[i] (flags) */
__global__ __launch_bounds__(kBlock) void k_or_flags(Cnt c_n, const uint8_t* a,
                                                     const uint8_t* b,
                                                     uint8_t* out) {
  const int64_t n = c_n.get();
  const int64_t i = gtid();
  if (i >= n) return;
  out[i] = (a[i] | b[i]) != 0;
}
void launch_or_flags(Cnt n, const uint8_t* a, const uint8_t* b, uint8_t* out,
                     PipeLaunch L) {
  if (n.cap <= 0) return;
  hipLaunchKernelGGL(k_or_flags, dim3((unsigned)((n.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, n, a, b, out);
}
/* out[pos[i]] = val[i] != 0 for the flagged i */
/* faces of the final tets -> tri_in */
__global__ __launch_bounds__(kBlock) void k_mark_faces(
    Cnt c_nk, const uint4* ftets, const uint64_t* erow, const uint32_t* ecol,
    const uint64_t* trow, const uint32_t* tk, uint8_t* tri_in) {
  const int64_t nk = c_nk.get();
  int64_t id = gtid();
  if (id >= nk) return;
  /* t = (i, j, k, l), i < j < k < l: faces ijk and ijl in the row of
   * edge ij, ikl in that of ik, jkl in that of jk; edges ij and ik are in
   * row i, so two row scans find the three edges */
  const uint4 t = ftets[id];
  int64_t eij, eik, f, g;
  find_row2(ecol, erow[t.x], erow[t.x + 1], t.y, t.z, eij, eik);
  const int64_t ejk = find_row(ecol, erow[t.y], erow[t.y + 1], t.z);
  if (eij >= 0) {
    find_row2(tk, trow[eij], trow[eij + 1], t.z, t.w, f, g);
    if (f >= 0) tri_in[f] = 1;
    if (g >= 0) tri_in[g] = 1;
  }
  if (eik >= 0 && (f = find_row(tk, trow[eik], trow[eik + 1], t.w)) >= 0)
    tri_in[f] = 1;
  if (ejk >= 0 && (f = find_row(tk, trow[ejk], trow[ejk + 1], t.w)) >= 0)
    tri_in[f] = 1;
}
void launch_mark_faces(Cnt nk, const uint4* ftets, const uint64_t* erow,
                       const uint32_t* ecol, const uint64_t* trow,
                       const uint32_t* tk, uint8_t* tri_in, PipeLaunch L) {
  if (nk.cap <= 0) return;
  hipLaunchKernelGGL(k_mark_faces, dim3((unsigned)((nk.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nk, ftets, erow, ecol, trow, tk,
                     tri_in);
}

/* edges of the final triangles -> edge_in (edge ids recorded at
 * enumeration: (i,j) = tedge, (i,k) = teik, (j,k) = tejk) */
__global__ __launch_bounds__(kBlock) void k_mark_edges(
    Cnt c_nt, const uint32_t* ftri, const uint32_t* tedge,
    const uint32_t* teik, const uint32_t* tejk, uint8_t* edge_in) {
  const int64_t nt = c_nt.get();
  int64_t id = gtid();
  if (id >= nt) return;
  uint32_t t = ftri[id];
  edge_in[tedge[t]] = 1;
  edge_in[teik[t]] = 1;
  edge_in[tejk[t]] = 1;
}
void launch_mark_edges(Cnt nt, const uint32_t* ftri, const uint32_t* tedge,
                       const uint32_t* teik, const uint32_t* tejk,
                       uint8_t* edge_in, PipeLaunch L) {
  if (nt.cap <= 0) return;
  hipLaunchKernelGGL(k_mark_edges, dim3((unsigned)((nt.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, ftri, tedge, teik, tejk,
                     edge_in);
}

/* ---- emission in original ids ---- */
__global__ __launch_bounds__(kBlock) void k_emit_tets(Cnt c_nk,
                                                      const uint4* ftets,
                                                      const uint32_t* perm,
                                                      int4* out) {
  const int64_t nk = c_nk.get();
  int64_t id = gtid();
  if (id >= nk) return;
  uint4 t = ftets[id];
  out[id] = make_int4((int)perm[t.x], (int)perm[t.y], (int)perm[t.z],
                      (int)perm[t.w]);
}
void launch_emit_tets(Cnt nk, const uint4* ftets, const uint32_t* perm,
                      int32_t* out, PipeLaunch L) {
  if (nk.cap <= 0) return;
  hipLaunchKernelGGL(k_emit_tets, dim3((unsigned)((nk.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nk, ftets, perm, (int4*)out);
}
__global__ __launch_bounds__(kBlock) void k_emit_tris(
    Cnt c_nt, const uint32_t* ftri, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* tk, const uint32_t* tedge,
    const uint32_t* perm, int32_t* out) {
  const int64_t nt = c_nt.get();
  int64_t id = gtid();
  if (id >= nt) return;
  uint32_t t = ftri[id];
  uint32_t e = tedge[t];
  uint64_t w = (uint64_t)id * 3;
  out[w] = (int32_t)perm[eleft[e]];
  out[w + 1] = (int32_t)perm[ecol[e]];
  out[w + 2] = (int32_t)perm[tk[t]];
}
void launch_emit_tris(Cnt nt, const uint32_t* ftri, const uint32_t* ecol,
                      const uint32_t* eleft, const uint32_t* tk,
                      const uint32_t* tedge, const uint32_t* perm,
                      int32_t* out, PipeLaunch L) {
  if (nt.cap <= 0) return;
  hipLaunchKernelGGL(k_emit_tris, dim3((unsigned)((nt.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, ftri, ecol, eleft, tk, tedge,
                     perm, out);
}

}  // namespace pac
