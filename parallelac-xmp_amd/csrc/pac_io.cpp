/*
 * pac_io.cpp -- CRD input, reference-format output and the deterministic
 * synthetic ball generators (host C++, no device code).
 *
 *   pac_read_crd      AlphaComplexParallel.cu:1069-1088 (token-based, so tab
 *                     separators and a trailing "#TER" line are accepted;
 *                     a short file is an error instead of garbage)
 *   pac_write_output  AlphaComplexParallel.cu:1834-1856 (byte-compatible)
 *   pac_write_crd     the layout of the reference's data files (the .crd files of data/)
 *   pac_read_pdb      PDB -> balls; no reference counterpart (the
 *                     reference's CRD files were made from PDB entries,
 *                     README.md:44-46, by a tool it does not ship)
 */
#include <cctype>
#include <cerrno>
#include <cmath>
#include <map>
#include <set>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pac.h"

extern "C" const char* pac_last_error(void);
void pac_set_error(const char* m);

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * 0x1p-53; }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 < 0x1p-60) u1 = 0x1p-60;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

uint64_t mix_seed(uint64_t a, uint64_t b) {
  SplitMix64 m(a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull));
  return m.next();
}

float draw_radius(SplitMix64& r) {
  double u = r.uni();
  if (u < 0.63) return 1.70f;
  if (u < 0.80) return 1.55f;
  if (u < 0.99) return 1.52f;
  return 1.80f;
}

/* Dart throwing with a hash grid of cell size min_sep.  `reject` can veto
 * a position (used to keep the lattice block free). */
template <class Reject>
bool poisson_sphere(int64_t n, double R, double min_sep, uint64_t seed,
                    std::vector<double>& pts, Reject reject) {
  pts.clear();
  if (n <= 0) return true;
  pts.reserve((size_t)n * 3);
  double cell = min_sep > 0 ? min_sep : 1.0;
  int64_t G = (int64_t)std::ceil(2 * R / cell) + 2;
  if (G > 4096) return false;
  std::vector<int32_t> head((size_t)(G * G * G), -1);
  std::vector<int32_t> nxt;
  nxt.reserve((size_t)n);
  SplitMix64 rng(seed);
  double ms2 = min_sep * min_sep;
  int64_t attempts = 0, limit = 2000 * n + 100000;
  auto cidx = [&](double v) {
    int64_t c = (int64_t)std::floor((v + R) / cell);
    return c < 0 ? 0 : (c >= G ? G - 1 : c);
  };
  while ((int64_t)nxt.size() < n) {
    if (++attempts > limit) return false;
    double x = (2 * rng.uni() - 1) * R, y = (2 * rng.uni() - 1) * R,
           z = (2 * rng.uni() - 1) * R;
    if (x * x + y * y + z * z > R * R) continue;
    if (reject(x, y, z)) continue;
    int64_t cx = cidx(x), cy = cidx(y), cz = cidx(z);
    bool ok = true;
    for (int64_t dz = -1; dz <= 1 && ok; ++dz)
      for (int64_t dy = -1; dy <= 1 && ok; ++dy)
        for (int64_t dx = -1; dx <= 1 && ok; ++dx) {
          int64_t a = cx + dx, b = cy + dy, c = cz + dz;
          if (a < 0 || b < 0 || c < 0 || a >= G || b >= G || c >= G) continue;
          for (int32_t p = head[(size_t)((c * G + b) * G + a)]; p >= 0;
               p = nxt[p]) {
            double ex = pts[3 * p] - x, ey = pts[3 * p + 1] - y,
                   ez = pts[3 * p + 2] - z;
            if (ex * ex + ey * ey + ez * ez < ms2) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    int32_t id = (int32_t)nxt.size();
    size_t h = (size_t)((cz * G + cy) * G + cx);
    nxt.push_back(head[h]);
    head[h] = id;
    pts.push_back(x);
    pts.push_back(y);
    pts.push_back(z);
  }
  return true;
}

double round_dec(double v, int decimals) {
  if (decimals < 0) return v;
  double sc = std::pow(10.0, decimals);
  return std::round(v * sc) / sc;
}

/* Radii by element.  C, N, O, S: the values of the reference's data files
 * (data/1grm.crd, data/1k4c.crd hold exactly the radii 1.548, 1.400,
 * 1.348, 1.808 in the proportions of carbon, nitrogen, oxygen and sulphur
 * of those proteins); the others: Bondi's van der Waals radii. */
const std::map<std::string, float>& default_radii() {
  static const std::map<std::string, float> t = {
      {"C", 1.548f}, {"N", 1.400f}, {"O", 1.348f}, {"S", 1.808f},
      {"H", 1.200f}, {"D", 1.200f}, {"P", 1.800f}, {"SE", 1.900f},
      {"F", 1.470f}, {"CL", 1.750f}, {"BR", 1.850f}, {"I", 1.980f}};
  return t;
}

std::string trim(const std::string& x) {
  size_t a = 0, b = x.size();
  while (a < b && std::isspace((unsigned char)x[a])) ++a;
  while (b > a && std::isspace((unsigned char)x[b - 1])) --b;
  return x.substr(a, b - a);
}

std::string upper(std::string x) {
  for (char& c : x) c = (char)std::toupper((unsigned char)c);
  return x;
}

/* columns [a, b) of a record (1-based a as in the PDB format) */
std::string field(const std::string& line, size_t a1, size_t b1) {
  const size_t a = a1 - 1, b = b1;
  if (line.size() <= a) return std::string();
  return line.substr(a, std::min(b, line.size()) - a);
}

bool parse_coord(const std::string& f, double* v) {
  const std::string t = trim(f);
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  *v = std::strtod(t.c_str(), &end);
  return end && *end == 0 && errno == 0 && std::isfinite(*v);
}

/* element of an atom record: columns 77-78, else the letters of the atom
 * name's columns 13-14 (the element is right-justified there) */
std::string element_of(const std::string& line) {
  std::string e = upper(trim(field(line, 77, 78)));
  if (!e.empty()) return e;
  std::string nm = field(line, 13, 14), r;
  for (char c : nm)
    if (std::isalpha((unsigned char)c)) r += (char)std::toupper((unsigned char)c);
  return r;
}

}  // namespace

extern "C" {

void pac_free(void* p) { std::free(p); }

int pac_read_crd(const char* path, float** xyzr, int64_t* n) {
  if (!path || !xyzr || !n) {
    pac_set_error("pac_read_crd: NULL argument");
    return PAC_ERR_INVALID;
  }
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    pac_set_error((std::string("cannot open ") + path).c_str());
    return PAC_ERR_IO;
  }
  std::string text;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  const char* p = text.c_str();
  char* end = nullptr;
  errno = 0;
  long long cnt = std::strtoll(p, &end, 10);
  if (end == p || errno || cnt < 0) {
    pac_set_error("CRD: missing ball count");
    return PAC_ERR_IO;
  }
  p = end;
  std::vector<float> v;
  v.reserve((size_t)cnt * 4);
  for (long long i = 0; i < cnt * 4; ++i) {
    errno = 0;
    float x = std::strtof(p, &end);
    if (end == p) {
      pac_set_error("CRD: short read (fewer than 4*N numbers)");
      return PAC_ERR_IO;
    }
    v.push_back(x);
    p = end;
  }
  float* out = (float*)std::malloc(std::max<size_t>(v.size(), 1) * sizeof(float));
  if (!out) return PAC_ERR_NOMEM;
  if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(float));
  *xyzr = out;
  *n = cnt;
  return PAC_OK;
}

int pac_write_crd(const char* path, const float* xyzr, int64_t n) {
  if (!path || n < 0 || (n > 0 && !xyzr)) {
    pac_set_error("pac_write_crd: bad argument");
    return PAC_ERR_INVALID;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    pac_set_error((std::string("cannot write ") + path).c_str());
    return PAC_ERR_IO;
  }
  std::fprintf(f, "%-13lld\n\n", (long long)n);
  for (int64_t i = 0; i < n; ++i)
    std::fprintf(f, "%.4f \t%.4f \t%.4f \t%.4f\n", (double)xyzr[4 * i],
                 (double)xyzr[4 * i + 1], (double)xyzr[4 * i + 2],
                 (double)xyzr[4 * i + 3]);
  std::fprintf(f, "#TER\n");
  if (std::fclose(f) != 0) {
    pac_set_error("CRD: write failed");
    return PAC_ERR_IO;
  }
  return PAC_OK;
}

int pac_read_pdb(const char* path, int flags, const char* radii, float** xyzr,
                 int64_t* n) {
  if (!path || !xyzr || !n) {
    pac_set_error("pac_read_pdb: NULL argument");
    return PAC_ERR_INVALID;
  }
  std::map<std::string, float> table = default_radii();
  float wildcard = -1.0f;
  if (radii) { /* "C=1.7,N=1.55,*=1.8" */
    std::string spec = radii;
    size_t at = 0;
    while (at <= spec.size()) {
      size_t comma = spec.find(',', at);
      if (comma == std::string::npos) comma = spec.size();
      const std::string item = trim(spec.substr(at, comma - at));
      at = comma + 1;
      if (item.empty()) continue;
      const size_t eq = item.find('=');
      char* end = nullptr;
      const double r = eq == std::string::npos
                           ? -1
                           : std::strtod(item.c_str() + eq + 1, &end);
      if (eq == std::string::npos || eq == 0 || !end || *end != 0 || !(r > 0) ||
          !std::isfinite(r)) {
        pac_set_error(("PDB: bad radius entry '" + item + "'").c_str());
        return PAC_ERR_INVALID;
      }
      const std::string el = upper(trim(item.substr(0, eq)));
      if (el == "*")
        wildcard = (float)r;
      else
        table[el] = (float)r;
    }
  }
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    pac_set_error((std::string("cannot open ") + path).c_str());
    return PAC_ERR_IO;
  }
  std::vector<float> v;
  std::set<std::string> seen; /* atoms already taken (alternate locations) */
  char buf[512];
  std::string line;
  long long lineno = 0;
  bool models = false;
  int rc = PAC_OK;
  std::string err;
  while (std::fgets(buf, sizeof(buf), f)) {
    line = buf;
    while (line.size() && (line.back() == '\n' || line.back() == '\r'))
      line.pop_back();
    if (line.size() == sizeof(buf) - 1) { /* over-long line: drop the rest */
      int c;
      while ((c = std::fgetc(f)) != EOF && c != '\n') {
      }
    }
    ++lineno;
    const std::string rec = field(line, 1, 6);
    if (rec == "MODEL ") {
      models = true;
      continue;
    }
    if (rec == "ENDMDL" && models) break; /* first model only */
    if (rec == "END   " || rec == "END") break;
    const bool atom = rec == "ATOM  ", het = rec == "HETATM";
    if (!atom && !(het && (flags & PAC_PDB_HETATM))) continue;
    const std::string res = trim(field(line, 18, 20));
    if (!(flags & PAC_PDB_WATER) &&
        (res == "HOH" || res == "WAT" || res == "DOD" || res == "H2O"))
      continue;
    const std::string el = element_of(line);
    if ((el == "H" || el == "D") && !(flags & PAC_PDB_HYDROGENS)) continue;
    /* alternate locations: the first one met for each atom */
    const std::string key = field(line, 13, 16) + "|" + field(line, 18, 27);
    if (field(line, 17, 17) != " " && !field(line, 17, 17).empty()) {
      if (seen.count(key)) continue;
    }
    seen.insert(key);
    double x, y, z;
    if (!parse_coord(field(line, 31, 38), &x) ||
        !parse_coord(field(line, 39, 46), &y) ||
        !parse_coord(field(line, 47, 54), &z)) {
      err = "PDB line " + std::to_string(lineno) + ": bad coordinates";
      rc = PAC_ERR_IO;
      break;
    }
    float r = wildcard;
    auto it = table.find(el);
    if (it != table.end()) r = it->second;
    if (!(r > 0)) {
      err = "PDB line " + std::to_string(lineno) + ": no radius for element '" +
            el + "' (give one in the radii table)";
      rc = PAC_ERR_INVALID;
      break;
    }
    v.push_back((float)x);
    v.push_back((float)y);
    v.push_back((float)z);
    v.push_back(r);
  }
  std::fclose(f);
  if (rc) {
    pac_set_error(err.c_str());
    return rc;
  }
  float* out = (float*)std::malloc(std::max<size_t>(v.size(), 1) * sizeof(float));
  if (!out) return PAC_ERR_NOMEM;
  if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(float));
  *xyzr = out;
  *n = (int64_t)(v.size() / 4);
  return PAC_OK;
}

int pac_write_output(const char* path, const float* xyzr_raw, int64_t n,
                     float solvent, const pac_result* r) {
  if (!path || !r || (n > 0 && !xyzr_raw)) {
    pac_set_error("pac_write_output: NULL argument");
    return PAC_ERR_INVALID;
  }
  int64_t c[4];
  int rc = pac_result_counts(r, c);
  if (rc) return rc;
  std::vector<int32_t> e((size_t)c[1] * 2), t((size_t)c[2] * 3),
      k((size_t)c[3] * 4);
  if ((rc = pac_result_copy_edges(r, e.data()))) return rc;
  if ((rc = pac_result_copy_triangles(r, t.data()))) return rc;
  if ((rc = pac_result_copy_tets(r, k.data()))) return rc;
  FILE* f = std::fopen(path, "w");
  if (!f) {
    pac_set_error((std::string("cannot write ") + path).c_str());
    return PAC_ERR_IO;
  }
  std::fprintf(f, "%d %ld %ld %ld\n", (int)n, (long)c[1], (long)c[2],
               (long)c[3]);
  for (int64_t i = 0; i < n; ++i) {
    volatile float rr = xyzr_raw[4 * i + 3] + solvent;
    std::fprintf(f, "%9.5lf %9.5lf %9.5lf %9.5lf\n", (double)xyzr_raw[4 * i],
                 (double)xyzr_raw[4 * i + 1], (double)xyzr_raw[4 * i + 2],
                 (double)rr);
  }
  for (int64_t i = 0; i < c[1]; ++i)
    std::fprintf(f, "%9d %9d\n", e[2 * i], e[2 * i + 1]);
  for (int64_t i = 0; i < c[2]; ++i)
    std::fprintf(f, "%9d %9d %9d\n", t[3 * i], t[3 * i + 1], t[3 * i + 2]);
  for (int64_t i = 0; i < c[3]; ++i)
    std::fprintf(f, "%9d %9d %9d %9d\n", k[4 * i], k[4 * i + 1], k[4 * i + 2],
                 k[4 * i + 3]);
  if (std::fclose(f) != 0) return PAC_ERR_IO;
  return PAC_OK;
}

int pac_synth_globule(int64_t n, double R, double min_sep, uint64_t seed,
                      int decimals, float* out) {
  if (n < 0 || (n > 0 && !out) || !(R > 0)) return PAC_ERR_INVALID;
  std::vector<double> pts;
  if (!poisson_sphere(n, R, min_sep, seed, pts,
                      [](double, double, double) { return false; })) {
    pac_set_error("synth: could not place the balls (too dense?)");
    return PAC_ERR_INVALID;
  }
  SplitMix64 rr(mix_seed(seed, 0x52414449ull));
  for (int64_t i = 0; i < n; ++i) {
    out[4 * i] = (float)round_dec(pts[3 * i], decimals);
    out[4 * i + 1] = (float)round_dec(pts[3 * i + 1], decimals);
    out[4 * i + 2] = (float)round_dec(pts[3 * i + 2], decimals);
    out[4 * i + 3] = draw_radius(rr);
  }
  return PAC_OK;
}

int pac_synth_globule_f64(int64_t n, double R, double min_sep, uint64_t seed,
                          double sigma, double* out) {
  if (n < 0 || (n > 0 && !out) || !(R > 0)) return PAC_ERR_INVALID;
  std::vector<double> pts;
  if (!poisson_sphere(n, R, min_sep, seed, pts,
                      [](double, double, double) { return false; })) {
    pac_set_error("synth: could not place the balls (too dense?)");
    return PAC_ERR_INVALID;
  }
  SplitMix64 rr(mix_seed(seed, 0x52414449ull));
  SplitMix64 jit(mix_seed(seed, 0x4a495454ull));
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) out[4 * i + a] = pts[3 * i + a] + sigma * jit.normal();
    out[4 * i + 3] = (double)draw_radius(rr);
  }
  return PAC_OK;
}

int pac_synth_batch(int64_t n_mol, int64_t nmin, int64_t nmax, uint64_t seed,
                    int64_t* offsets, float* out) {
  if (n_mol < 0 || !offsets || nmin < 1 || nmax < nmin) return PAC_ERR_INVALID;
  SplitMix64 sz(seed);
  offsets[0] = 0;
  std::vector<int64_t> sizes((size_t)n_mol);
  double l0 = std::log((double)nmin), l1 = std::log((double)nmax);
  for (int64_t m = 0; m < n_mol; ++m) {
    int64_t k = (int64_t)std::llround(std::exp(l0 + (l1 - l0) * sz.uni()));
    k = std::max(nmin, std::min(nmax, k));
    sizes[(size_t)m] = k;
    offsets[m + 1] = offsets[m] + k;
  }
  if (!out) return PAC_OK;
  for (int64_t m = 0; m < n_mol; ++m) {
    int64_t k = sizes[(size_t)m];
    double R = std::cbrt(3.0 * (double)k / (4.0 * M_PI * 0.05));
    int rc = pac_synth_globule(k, R, 1.3, mix_seed(seed, (uint64_t)m + 1), 3,
                               out + 4 * offsets[m]);
    if (rc) return rc;
  }
  return PAC_OK;
}

int pac_synth_stress(int64_t n, double R, uint64_t seed, double frac,
                     double spacing, double lattice_r, float* out) {
  if (n < 0 || (n > 0 && !out) || !(R > 0) || frac < 0 || frac > 1 ||
      !(spacing > 0))
    return PAC_ERR_INVALID;
  int64_t m = (int64_t)std::llround(std::cbrt(frac * (double)n));
  int64_t nl = m * m * m;
  if (nl > n) {
    m -= 1;
    nl = m * m * m;
  }
  double half = 0.5 * (double)(m - 1) * spacing;
  double guard = half + 0.65;
  std::vector<double> pts;
  if (!poisson_sphere(n - nl, R, 1.3, seed, pts,
                      [&](double x, double y, double z) {
                        return std::fabs(x) <= guard && std::fabs(y) <= guard &&
                               std::fabs(z) <= guard;
                      })) {
    pac_set_error("synth: could not place the balls (too dense?)");
    return PAC_ERR_INVALID;
  }
  SplitMix64 rr(mix_seed(seed, 0x52414449ull));
  int64_t w = 0;
  for (int64_t i = 0; i < n - nl; ++i, ++w) {
    out[4 * w] = (float)round_dec(pts[3 * i], 3);
    out[4 * w + 1] = (float)round_dec(pts[3 * i + 1], 3);
    out[4 * w + 2] = (float)round_dec(pts[3 * i + 2], 3);
    out[4 * w + 3] = draw_radius(rr);
  }
  for (int64_t a = 0; a < m; ++a)
    for (int64_t b = 0; b < m; ++b)
      for (int64_t c = 0; c < m; ++c, ++w) {
        out[4 * w] = (float)round_dec(-half + (double)c * spacing, 3);
        out[4 * w + 1] = (float)round_dec(-half + (double)b * spacing, 3);
        out[4 * w + 2] = (float)round_dec(-half + (double)a * spacing, 3);
        out[4 * w + 3] = (float)lattice_r;
      }
  return PAC_OK;
}

}  // extern "C"
