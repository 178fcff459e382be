/*
 * pac_cli.cpp -- the `parallelac` command, a thin wrapper over libpac.
 *
 *   parallelac <crd-file> <out-file> <sol-rad> <alpha> [--staged|--plain]
 *              [--reference-raw] [--skip-tet-ac2]
 *
 * Same positional interface, output file and stdout summary as the
 * reference (AlphaComplexParallel.cu:1053-1057, :1809-1856).  Invoked as
 * `parallelac-largeData ... <chunk-size>` (AlphaComplexLargeData.cu:
 * 1054-1058) the input is processed in chunks of 512 * 2^k >= chunk-size
 * cell-sorted balls (pac_options.chunk_balls) and the per-chunk progress
 * lines of that driver are printed (:1235); the lists are those of one pass.
 */
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pac.h"

int main(int argc, char** argv) {
  std::string prog = argv[0];
  bool large = prog.find("largeData") != std::string::npos;
  int npos = large ? 6 : 5;
  if (argc < npos) {
    if (large)
      std::printf("Usage : parallelac-largeData <crd-file> <out-file> <sol-rad> <alpha> <chunk-size>\n");
    else
      std::printf("Usage : parallelac <crd-file> <out-file> <sol-rad> <alpha>\n");
    return 1;
  }
  pac_options opt;
  pac_options_init(&opt);
  if (large) opt.chunk_balls = std::max(1, std::atoi(argv[5]));
  for (int i = npos; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--staged")) opt.precision = PAC_PREC_STAGED;
    else if (!std::strcmp(argv[i], "--plain")) opt.precision = PAC_PREC_PLAIN_FP64;
    else if (!std::strcmp(argv[i], "--reference-raw")) opt.cell_policy = PAC_CELL_REFERENCE_RAW;
    else if (!std::strcmp(argv[i], "--skip-tet-ac2")) opt.skip_tet_ac2 = 1;
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  float solvent = (float)std::atof(argv[3]);
  float alpha = (float)std::atof(argv[4]);
  float* xyzr = nullptr;
  int64_t n = 0;
  int rc = pac_read_crd(argv[1], &xyzr, &n);
  if (rc) {
    std::fprintf(stderr, "error: %s\n", pac_last_error());
    return 3;
  }
  std::printf("Successfully read the file %s ...\n", argv[1]);
  std::printf("The given file has %d atoms. \n\nStarting alpha complex computation for "
              "alpha = %.3f and solvent radius = %.3f ...\n\n",
              (int)n, alpha, solvent);
  pac_result* res = nullptr;
  rc = pac_compute(xyzr, n, solvent, alpha, &opt, &res);
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, pac_last_error());
    pac_free(xyzr);
    return 4;
  }
  if (large) {
    int64_t chunk = 512;
    while (chunk < opt.chunk_balls) chunk *= 2;
    for (int64_t at = 0; at < n; at += chunk)
      std::printf("Processed %d atoms ... \n", (int)at);
  }
  int64_t c[4];
  pac_result_counts(res, c);
  pac_stats st;
  pac_result_stats(res, &st);
  std::printf("Alpha complex computed.\n\n");
  std::printf("Number of vertices: %15d\n", (int)n);
  std::printf("Number of edges: %18ld\n", (long)c[1]);
  std::printf("Number of triangles: %14ld\n", (long)c[2]);
  std::printf("Number of tetrahedra: %13ld\n", (long)c[3]);
  std::printf("Total simplices: %18ld\n", (long)(n + c[1] + c[2] + c[3]));
  const double* t = st.phase_ms;
  std::printf("\nTime taken (ms):\nData transfer: %20.3f\nGrid computation: %17.3f\nPotential edges: %18.3f\n"
              "Potential triangles: %14.3f\nPotential tetrahedra: %13.3f\nAC2 test for tetrahedra: %10.3f\n"
              "AC2 test for triangles: %11.3f\nAC2 test for edges: %15.3f\n",
              t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7]);
  double tot = 0;
  for (int i = 0; i < 8; ++i) tot += t[i];
  std::printf("Total time taken (ms): %12.3f\n", tot);
  std::printf("\nWriting computed alpha complex to the file %s ...\n", argv[2]);
  rc = pac_write_output(argv[2], xyzr, n, solvent, res);
  pac_result_free(res);
  pac_free(xyzr);
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, pac_last_error());
    return 5;
  }
  return 0;
}
