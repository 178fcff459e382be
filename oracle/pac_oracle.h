/*
 * pac_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's weighted alpha complex pipeline
 * (forest113/ParallelAC-xmp: AlphaComplexLargeData.cu:65-1011 kernel math,
 * AlphaComplexParallel.cu:1053-1856 single-pass orchestration and closure),
 * with every sign decision made exactly.  It is the checker for the HIP
 * library and the CPU baseline that bench.py times beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Semantics (identical to the product's, so outputs must match bit for bit):
 *   - balls r = fl(r_raw + solvent) in the input float type, weights w = r^2
 *     taken exactly, alpha exact;
 *   - sorted ids: stable sort by the reference's FP32 cell key
 *     (AlphaComplexParallel.cu:107-120, :1120-1124, :1157);
 *   - alpha edges / triangles / tetrahedra: rho <= alpha of the smallest
 *     orthosphere, degenerate (coincident / collinear / coplanar) rejected;
 *   - tet emptiness: insphere x orient rule of AlphaComplexParallel.cu:653-683
 *     with Edelsbrunner-Muecke SoS on exact ties (original index order);
 *   - triangle / edge emptiness: reject when some other ball has power
 *     <= rho (AlphaComplexParallel.cu:838, :1005);
 *   - closure as AlphaComplexParallel.cu:1490-1788.
 */
#ifndef PAC_ORACLE_H
#define PAC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cell_policy: 0 = exact (stencil widened until grid-independent),
 *              1 = reference_raw (27-cell stencil on the reference grid). */
typedef struct pac_oracle_result pac_oracle_result;

int pac_oracle_run_f32(const float* xyzr_raw, int64_t n, float solvent,
                       float alpha, int skip_tet_ac2, int cell_policy,
                       int nthreads, pac_oracle_result** out);
int pac_oracle_run_f64(const double* xyzr_raw, int64_t n, double solvent,
                       double alpha, int skip_tet_ac2, int cell_policy,
                       int nthreads, pac_oracle_result** out);

/* dim 1 = edges, 2 = triangles, 3 = tetrahedra; ids are original 0-based,
 * rows in the reference's order. */
int64_t pac_oracle_count(const pac_oracle_result* r, int dim);
void pac_oracle_copy(const pac_oracle_result* r, int dim, int32_t* dst);
/* sorted -> original permutation (n entries) */
void pac_oracle_perm(const pac_oracle_result* r, int32_t* dst);
/* stats: [0] n, [1] E_cand, [2] E_alpha, [3] T_cand, [4] T_alpha,
 * [5] K_cand, [6] K_alpha, [7] K_final, [8] T_final, [9] E_final,
 * [10] exact filter fallbacks, [11] exact AC2 fallbacks, [12] SoS calls,
 * [13] AC2 pair tests, [14] k_enum, [15] k_ac2, [16] scale S */
void pac_oracle_stats(const pac_oracle_result* r, int64_t* out);
/* 8 phase times (ms) with the reference's labels, in order */
void pac_oracle_times(const pac_oracle_result* r, double* out);
void pac_oracle_free(pac_oracle_result* r);

/* Single predicates on balls given as (x, y, z, r) with r already including
 * the solvent.  Return 1 keep / 0 reject. */
int pac_oracle_edge_keep(const float* b0, const float* b1, float alpha);
int pac_oracle_tri_keep(const float* b0, const float* b1, const float* b2,
                        float alpha);
int pac_oracle_tet_keep(const float* b0, const float* b1, const float* b2,
                        const float* b3, float alpha);
/* 1 if ball e violates the smallest orthosphere of the triangle / edge */
int pac_oracle_tri_violated(const float* b0, const float* b1, const float* b2,
                            const float* e);
int pac_oracle_edge_violated(const float* b0, const float* b1, const float* e);
/* 1 if e violates tet (b0..b3); ids are the SoS ranks (original ids) */
int pac_oracle_tet_violated(const float* const* b4, const float* e,
                            const int64_t* ids5);
/* sign of Edelsbrunner-Muecke SoS of det[x y z |p|^2-w 1] over 5 balls in
 * argument order; ids give the perturbation order. */
int pac_oracle_sos_det5(const float* const* b5, const int64_t* ids5);
/* exact sign of det[x y z |p|^2-w 1] (no perturbation) */
int pac_oracle_det5(const float* const* b5);
/* exact sign of orient3d (det[a-d; b-d; c-d]) */
int pac_oracle_orient3d(const float* a, const float* b, const float* c,
                        const float* d);
/* SoS term table: writes up to cap entries of (rows mask, cols mask, sign),
 * returns the number of terms. */
int pac_oracle_sos_table(int32_t* out3, int cap);

#ifdef __cplusplus
}
#endif
#endif
