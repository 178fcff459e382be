/*
 * pac.h -- C ABI of the MI355X weighted alpha complex engine (libpac.so).
 *
 * Drop-in boundary.  The reference (forest113/ParallelAC-xmp) has no
 * library: its only interface is the process CLI
 *     parallelac <crd-file> <out-file> <sol-rad> <alpha>
 * (AlphaComplexParallel.cu:1053-1057; chunked variant
 * AlphaComplexLargeData.cu:1054-1058).  Each entry point below replaces one
 * step of that program; the CLI `parallelac` built from this library is a
 * thin wrapper over them and writes the reference's file format.
 *
 * Conventions: every function returns PAC_OK (0) or a negative status and
 * never prints; pac_last_error() gives a thread-local message.  Results are
 * library-owned and opaque; free them with pac_result_free().  Distinct
 * results / streams may be used from distinct threads; one result is not
 * re-entrant.  Simplex vertex ids are original 0-based input positions.
 */
#ifndef PAC_H
#define PAC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PAC_OK = 0,
  PAC_ERR_IO = -1,          /* file missing / unreadable / short read */
  PAC_ERR_INVALID = -2,     /* bad argument, non-finite input, empty grid */
  PAC_ERR_DEGENERATE = -3,  /* reserved: degenerate input reported */
  PAC_ERR_NOMEM = -4,       /* host or device allocation failed */
  PAC_ERR_DEVICE = -5,      /* HIP runtime error */
  PAC_ERR_OVERFLOW = -6,    /* a count would exceed the index width */
  PAC_ERR_RANGE = -7,       /* input exponent range exceeds exact capacity */
  PAC_ERR_STALE = -8        /* a recorded graph no longer fits its input */
};

enum { PAC_CELL_EXACT = 0, PAC_CELL_REFERENCE_RAW = 1 };
enum { PAC_PREC_PLAIN_FP64 = 0, PAC_PREC_STAGED = 1 };
enum { PAC_ORDER_REFERENCE = 0, PAC_ORDER_CANONICAL = 1 };

typedef struct pac_options {
  int32_t cell_policy;   /* PAC_CELL_*: exact = grid-independent stencil  */
  int32_t precision;     /* PAC_PREC_*: FP64->int, or FP32->FP64->DD->int */
  int32_t skip_tet_ac2;  /* debug / fixture mode: keep all alpha tets     */
  int32_t order;         /* PAC_ORDER_*: row order of the output lists    */
  int32_t device_inputs; /* 1: xyzr is a device pointer (e.g. torch)      */
  int32_t device;        /* HIP device ordinal                            */
  void* stream;          /* hipStream_t to run on, NULL = library stream  */
  int32_t block;         /* threads per block of the persistent second
                            emptiness phase: 64, 128 or 256 (0 = 256)     */
  int32_t profile;       /* 1: time the main kernels with hipEvents       */
  int32_t ac2_budget;    /* emptiness phase-one pair budget: > 0 splits
                            each scan after that many balls, the rest
                            finished by the persistent second phase
                            (0 = 4); < 0: one phase, every scan to its end */
  int32_t batch_balls;   /* batch: max balls per pipeline pass (0 = 4Mi;
                            a recorded batch, pac_graph_create: 2Mi)      */
  int32_t ac2_grid;      /* blocks of the second emptiness phase (0=2048) */
  int32_t ac2_fine;      /* exact policy: emptiness scans over the grid's
                            cells split in ac2_fine (2..8) along x
                            (0 = 2, < 0 or 1: the grid's own cells)       */
  int32_t ac2_quick;     /* fine scan grid: 1 = a quick rejection pass over
                            the alpha tets first (0 = off)                */
  int32_t chunk_balls;   /* > 0: the LD driver (AlphaComplexLargeData.cu:
                            1227-1233): the cell-sorted balls are processed
                            in chunks of 512 * 2^k >= chunk_balls, each by
                            its own pipeline run over the chunk and a halo;
                            the lists equal one pass.  0 = one pass.
                            pac_compute / pac_compute_f64 only.           */
  int32_t filtration;    /* 1: also decide, for every output edge and
                            triangle, whether it enters the filtration at
                            its own rho (pac_result_copy_birth /
                            pac_result_copy_filtration).  Reference row
                            order, with the tet emptiness test;
                            pac_compute / pac_compute_f64 only.           */
  int32_t ac2_chunk;     /* entries of the left-over list each wave of the
                            persistent second emptiness phase takes at a
                            time, wave w chunks w, w + W, ... (rounded up
                            to a multiple of 64; 0 = 64)                  */
} pac_options;

typedef struct pac_result pac_result;

/* Per-run statistics.  stage_evals[p][s]: predicate family p (0 edge filter,
 * 1 triangle filter, 2 tet filter, 3 tet emptiness, 4 triangle emptiness,
 * 5 edge emptiness) evaluated at stage s (0 FP32, 1 FP64, 2 double-double,
 * 3 wide integer); sos_calls counts exact ties broken by SoS. */
typedef struct pac_stats {
  int64_t n;
  int64_t edge_candidates, edge_alpha, edges_final;
  int64_t tri_candidates, tri_alpha, tris_final;
  int64_t tet_candidates, tet_alpha, tets_final;
  int64_t stage_evals[6][4];
  int64_t sos_calls;
  int64_t ac2_pairs[3]; /* (simplex, ball) pairs tested: tet, tri, edge */
  int32_t k_enum, k_ac2, scale_bits, grid_cells;
  /* phase times in ms with the reference's labels
   * (AlphaComplexParallel.cu:1818-1830): Data transfer, Grid computation,
   * Potential edges, Potential triangles, Potential tetrahedra,
   * AC2 test for tetrahedra, AC2 test for triangles, AC2 test for edges */
  double phase_ms[8];
  double total_ms;
  /* with pac_options.profile: device time (ms) of the main kernels, in
   * order edge_fill, tri_fill, tet_fill, tet_ac2, tri_ac2, edge_ac2,
   * radix_sort, tet_ac2_exact (hipEvents on the run's stream) */
  double kernel_ms[8];
} pac_stats;

void pac_options_init(pac_options* opt);

/* Whole pipeline on one GPU.  xyzr: n x (x, y, z, r_raw) FP32 in input
 * order; balls carry fl32(r_raw + solvent) (AlphaComplexParallel.cu:1086).
 * Replaces AlphaComplexParallel.cu:1126-1806 (everything between parsing
 * and writing). */
int pac_compute(const float* xyzr, int64_t n, float solvent, float alpha,
                const pac_options* opt, pac_result** out);

/* Same with FP64 inputs (continuous coordinates; no reference counterpart,
 * the reference is FP32-only: predicates.h:142). */
int pac_compute_f64(const double* xyzr, int64_t n, double solvent,
                    double alpha, const pac_options* opt, pac_result** out);

/* Allocation-free pac_compute.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops and harnesses
 * that forbid hipMalloc).  Every intermediate buffer comes from a device
 * scratch array of PAC_NOALLOC_SCRATCH_BYTES compiled into the library
 * (one per GPU, held until the process ends); the call makes no hipMalloc,
 * hipFree, pinned-memory or event allocation.  xyzr: n x (x, y, z, r_raw)
 * FP32, device memory with opt->device_inputs, else host memory; n <=
 * PAC_NOALLOC_MAX_BALLS, and an input whose intermediates outgrow the
 * scratch fails with PAC_ERR_NOMEM (8192 balls of protein density fit).
 * The lists are written to edges / triangles / tets (device or host memory
 * of cap_* rows of 2 / 3 / 4 int32, the same rows pac_compute returns in
 * opt->order) and counts[4] (host) receives vertices, edges, triangles,
 * tets; if a list exceeds its capacity nothing is written to the lists,
 * counts[] holds the sizes needed and PAC_ERR_OVERFLOW is returned.
 * opt->chunk_balls, opt->filtration and opt->profile must be 0.  Calls on
 * one GPU share its scratch: this library serializes them within a process
 * and drains the stream before returning; calls that overlap on one GPU
 * through any other route race.  Not capturable in a HIP graph (it reads
 * counts back to size its stages). */
#define PAC_NOALLOC_SCRATCH_BYTES (256u << 20)
#define PAC_NOALLOC_MAX_BALLS 8192
int pac_compute_noalloc(const float* xyzr, int64_t n, float solvent,
                        float alpha, const pac_options* opt, int32_t* edges,
                        int64_t cap_edges, int32_t* triangles,
                        int64_t cap_triangles, int32_t* tets, int64_t cap_tets,
                        int64_t counts[4]);

/* counts[0..3] = vertices, edges, triangles, tetrahedra
 * (AlphaComplexParallel.cu:1809-1816). */
int pac_result_counts(const pac_result* r, int64_t counts[4]);
/* Copy simplex lists (row-major, 2/3/4 int32 per row) to host or device
 * memory (AlphaComplexParallel.cu:1794-1806 + the id mapping at
 * :1842-1856). */
int pac_result_copy_edges(const pac_result* r, int32_t* dst);
int pac_result_copy_triangles(const pac_result* r, int32_t* dst);
int pac_result_copy_tets(const pac_result* r, int32_t* dst);
/* Device pointer of the list for dim 1/2/3 (valid until pac_result_free). */
const int32_t* pac_result_device_list(const pac_result* r, int dim);
/* FP64 alpha values (squared orthoradius rho) of the output simplices of
 * dim 1/2/3, in list order, to host or device memory.  The reference computes them for triangles and
 * tets (float4 .w, AlphaComplexParallel.cu:370, :603) and discards them. */
int pac_result_copy_alpha(const pac_result* r, int dim, double* rho);
int pac_result_stats(const pac_result* r, pac_stats* st);
void pac_result_free(pac_result* r);

/* Filtration of the complex (a result computed with opt.filtration = 1).
 * The value at which a simplex enters the filtration of weighted alpha
 * complexes: a ball at -r^2; an edge or triangle whose smallest
 * orthosphere is empty (no other ball of power <= rho: the reference's
 * emptiness rule, AlphaComplexLargeData.cu:838, :971) at its own rho; any
 * other edge or triangle at the smallest value of the triangles / tets of
 * the complex it is a face of; a tet at its rho.  The complex at any
 * alpha' <= alpha is exactly the simplices with value <= alpha'.  FP64.
 * No reference counterpart: the reference computes rho for triangles and
 * tets (float4 .w, AlphaComplexParallel.cu:370, :603) and discards it.
 * copy_birth: values of the simplices of dim 0..3 (balls in input order,
 * the lists in list order).  copy_filtration: all n + E + T + K simplices
 * sorted by value, ties in the order balls, edges, triangles, tets (so
 * every simplex follows its faces), as (dim, index in its list, value). */
int pac_result_copy_birth(const pac_result* r, int dim, double* value);
int pac_result_copy_filtration(const pac_result* r, int32_t* dim,
                               int64_t* index, double* value);

/* Batched ensemble: n_mol molecules concatenated, offsets[n_mol+1]
 * (int64, offsets[0] = 0) into the ball array (n x 4 FP32, device memory
 * with opt.device_inputs).  Every molecule gets its own reference grid and
 * its lists equal those of pac_compute on that molecule alone; molecules
 * are packed into pipeline passes of at most opt.batch_balls balls, each
 * pass processing all its molecules with one launch per kernel.  No
 * reference counterpart (the reference handles one molecule per process). */
typedef struct pac_batch_result pac_batch_result;
int pac_compute_batch(const float* xyzr, const int64_t* offsets, int64_t n_mol,
                      float solvent, float alpha, const pac_options* opt,
                      pac_batch_result** out);
/* per-molecule counts: out[m*4 + {0..3}] */
int pac_batch_counts(const pac_batch_result* r, int64_t* out);
/* lists of molecule m (ids local to the molecule) */
int pac_batch_copy(const pac_batch_result* r, int64_t m, int dim, int32_t* dst);
int pac_batch_stats(const pac_batch_result* r, pac_stats* st);
void pac_batch_free(pac_batch_result* r);

/* Recorded batch: the whole pipeline of every pass of a batch captured
 * once as one HIP graph, so that a run is a single hipGraphLaunch with no
 * host round trip.  pac_graph_create runs the batch eagerly once (this
 * fixes each pass's grids and measures every count), records each pass
 * with buffers of those counts plus `headroom` (a fraction, e.g. 0.125;
 * < 0: the default 0.125) and checks one uncaptured run of the recording
 * against the eager run.  Each launch recomputes everything from the
 * input buffer: the graph's own copy of xyzr (replace its contents with
 * pac_graph_set_input) or, with opt.device_inputs, the caller's device
 * array itself.  pac_graph_result waits for the last launch and returns
 * its lists as a batch result (pac_batch_*), equal to pac_compute_batch's.
 * A launch whose input moved a molecule's bounds (corners, largest radius
 * or |coordinate|, finer coordinate lsb) reports PAC_ERR_STALE, one whose
 * counts outgrew the headroom PAC_ERR_OVERFLOW: record again.  The
 * reference row order only (opt.order).  No reference counterpart. */
typedef struct pac_graph pac_graph;
int pac_graph_create(const float* xyzr, const int64_t* offsets, int64_t n_mol,
                     float solvent, float alpha, const pac_options* opt,
                     double headroom, pac_graph** out);
/* replace the graph's copy of the input (host or device pointer, same
 * ball count), ordered after earlier launches on `stream` */
int pac_graph_set_input(pac_graph* g, const float* xyzr, void* stream);
/* one run of the whole batch on `stream` (NULL: the graph's stream) */
int pac_graph_launch(pac_graph* g, void* stream);
int pac_graph_result(pac_graph* g, pac_batch_result** out);
/* kernel nodes in the graph, graph device memory in bytes */
int pac_graph_info(const pac_graph* g, int64_t* nodes, int64_t* bytes);
void pac_graph_free(pac_graph* g);

/* Autotuner: times the pipeline on the given balls for each candidate
 * setting of the tunables and writes the fastest into opt->ac2_fine (the
 * emptiness scan grid, exact cell policy: x subdivision 2, 4 or the
 * reference grid, tried first), opt->ac2_quick (with a fine grid: the
 * quick rejection pass on or off), then opt->ac2_budget / opt->ac2_grid
 * (one phase, or a phase-one pair budget of 2, 4, 8 or 16 x second-phase
 * grid, over the chosen scan grid), then, with a second phase,
 * opt->block / opt->ac2_chunk (its block shape and queue granularity);
 * the other fields of *opt
 * are used as given.  offsets ==
 * NULL: the n balls are one molecule; otherwise they are a batch
 * (offsets[n_mol+1]) and the leading molecules, up to about 2^20 balls,
 * are the sample.  best_ms (may be NULL) receives the fastest wall time
 * of one run.  Outputs never depend on these settings.  No reference
 * counterpart. */
int pac_autotune(const float* xyzr, int64_t n, const int64_t* offsets,
                 int64_t n_mol, float solvent, float alpha, pac_options* opt,
                 double* best_ms);

/* Adjoint of the per-simplex alpha values with the combinatorics held
 * fixed: given dL/drho for the output triangles and tets (list order),
 * accumulate dL/d(x,y,z) (n x 3) and dL/d(r) (n) in input order.  No
 * reference counterpart. */
int pac_alpha_adjoint(const pac_result* r, const double* dl_drho_tri,
                      const double* dl_drho_tet, double* dl_dxyz,
                      double* dl_dr);

/* Device memory of the library's buffer pool on a device (0s before its
 * first use): bytes in live buffers, their high-water mark since the last
 * reset (reset_peak != 0 resets it to the current use), and bytes held
 * (live plus cached for reuse); sizes by power-of-two class.
 * pac_memory_trim returns the cached blocks to the device.  What the
 * chunked driver (opt.chunk_balls) bounds.  No reference counterpart. */
int pac_memory_stats(int device, int64_t* in_use, int64_t* peak,
                     int64_t* held, int reset_peak);
int pac_memory_trim(int device);

/* CRD input (AlphaComplexParallel.cu:1069-1088): returns a malloc'ed
 * n x 4 FP32 array (x, y, z, r_raw); free with pac_free(). */
int pac_read_crd(const char* path, float** xyzr, int64_t* n);
/* Output file in the reference's format (AlphaComplexParallel.cu:1834-1856). */
int pac_write_output(const char* path, const float* xyzr_raw, int64_t n,
                     float solvent, const pac_result* r);
void pac_free(void* p);
/* CRD file in the layout of the reference's data files (the .crd files of data/):
 * the count, a blank line, "x \ty \tz \tr" rows with 4 decimals, "#TER".
 * No reference counterpart (the reference only reads CRD). */
int pac_write_crd(const char* path, const float* xyzr, int64_t n);
/* PDB -> balls (x, y, z, r) for pac_compute / pac_write_crd.  The ATOM
 * records of the first MODEL (and the HETATM records with PAC_PDB_HETATM),
 * in file order; of an atom's alternate locations the first met; hydrogens
 * and waters (HOH, WAT, DOD, H2O) left out unless PAC_PDB_HYDROGENS /
 * PAC_PDB_WATER.  Radius by element (columns 77-78, else the atom name's
 * columns 13-14): C 1.548, N 1.400, O 1.348, S 1.808 -- the radii of the
 * reference's data files -- and Bondi radii for H, D, P, SE, F, CL, BR, I;
 * `radii` (may be NULL), e.g. "C=1.7,N=1.55,*=1.8", overrides entries and
 * "*" gives every other element a radius.  An element without a radius is
 * PAC_ERR_INVALID.  Returns a malloc'ed n x 4 array; free with pac_free().
 * No reference counterpart: the reference's CRD files were made from PDB
 * entries (README.md:44-46) by a tool it does not ship. */
enum { PAC_PDB_HETATM = 1, PAC_PDB_HYDROGENS = 2, PAC_PDB_WATER = 4 };
int pac_read_pdb(const char* path, int flags, const char* radii, float** xyzr,
                 int64_t* n);

/* Synthetic protein-like ball sets (deterministic; SplitMix64 seeds).
 * globule: Poisson-disk centres (min separation min_sep) in a sphere,
 * radii drawn from {1.70:.63, 1.55:.17, 1.52:.19, 1.80:.01}, coordinates
 * rounded to 10^-decimals (decimals < 0: no rounding). */
int pac_synth_globule(int64_t n, double sphere_radius, double min_sep,
                      uint64_t seed, int decimals, float* xyzr_out);
int pac_synth_globule_f64(int64_t n, double sphere_radius, double min_sep,
                          uint64_t seed, double jitter_sigma,
                          double* xyzr_out);
/* ensemble sizes: n_mol sizes log-uniform in [nmin, nmax]; offsets_out has
 * n_mol+1 entries; call with xyzr_out = NULL first to size the buffer. */
int pac_synth_batch(int64_t n_mol, int64_t nmin, int64_t nmax, uint64_t seed,
                    int64_t* offsets_out, float* xyzr_out);
/* stress: lattice_frac of the balls form a central cubic lattice of
 * spacing `spacing` with raw radius lattice_r, the rest Poisson-disk. */
int pac_synth_stress(int64_t n, double sphere_radius, uint64_t seed,
                     double lattice_frac, double spacing, double lattice_r,
                     float* xyzr_out);

const char* pac_last_error(void);
const char* pac_version(void);

#ifdef __cplusplus
}
#endif
#endif
