/*
 * pac_pdb2crd.cpp -- the `pdb2crd` command: a PDB entry to the CRD input
 * of `parallelac` (pac_read_pdb + pac_write_crd).
 *
 *   pdb2crd <pdb-file> <crd-file> [--hetatm] [--hydrogens] [--water]
 *           [--radii C=1.7,N=1.55,*=1.8]
 *
 * No reference counterpart: the reference's CRD files were made from PDB
 * entries (README.md:44-46) by a tool it does not ship.
 */
#include <cstdio>
#include <cstring>

#include "../../include/pac.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("Usage : pdb2crd <pdb-file> <crd-file> [--hetatm] [--hydrogens] "
                "[--water] [--radii E=r,...]\n");
    return 1;
  }
  int flags = 0;
  const char* radii = nullptr;
  for (int i = 3; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--hetatm")) flags |= PAC_PDB_HETATM;
    else if (!std::strcmp(argv[i], "--hydrogens")) flags |= PAC_PDB_HYDROGENS;
    else if (!std::strcmp(argv[i], "--water")) flags |= PAC_PDB_WATER;
    else if (!std::strcmp(argv[i], "--radii") && i + 1 < argc) radii = argv[++i];
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  float* xyzr = nullptr;
  int64_t n = 0;
  int rc = pac_read_pdb(argv[1], flags, radii, &xyzr, &n);
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, pac_last_error());
    return 3;
  }
  rc = pac_write_crd(argv[2], xyzr, n);
  pac_free(xyzr);
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, pac_last_error());
    return 4;
  }
  std::printf("Wrote %lld balls to %s\n", (long long)n, argv[2]);
  return 0;
}
