// Readers and writer for the reference's text formats (host only, no GPU calls).
//
//   position.txt     num, E1, nu1, then num x (x y z r)          configElastic3D.cpp:105-125
//   BEM.txt          title, E0, nu0, NN, NE (one per line), then NN x 3 nodes, NE x 4 one-based
//                    connectivity, 3NN x 2 (flag, value), NE x 24 traction flags/values
//                                                                 configElastic3D.cpp:32-103
//   postprocess.txt  nump, then nump x 3                          configElastic3D.h:113-126
//   stress.txt       nump rows "s00 s01 s02 s11 s12 s22", C++ std::scientific (6 digits)
//                                                                 postProcessorElastic3D.cpp:578-585
//
// The header lines of BEM.txt are line-oriented (the reference reads them with getline + stod);
// everything after them is whitespace-separated tokens. Carriage returns are tolerated anywhere.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibem3d.h"

struct ibem3d_inputs {
  ibem3d_problem prob{};
  std::string title;
  std::vector<double> nodes, bc, tbc, centres, radii, points;
  std::vector<int> conn;
};

namespace {

std::string g_io_error;

bool slurp(const std::string& path, std::string& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t k;
  out.clear();
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
  std::fclose(f);
  return true;
}

// Token cursor over a whole file.
struct Cursor {
  const char* p;
  const char* end;
  std::string file;
  bool ok = true;

  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v')) ++p;
  }
  double num() {
    skip_ws();
    if (p >= end) return fail("unexpected end of file");
    char* q = nullptr;
    errno = 0;
    const double v = std::strtod(p, &q);
    if (q == p) return fail("expected a number");
    p = q;
    return v;
  }
  long integer() {
    const double v = num();
    if (!ok) return 0;
    if (v != (double)(long)v) {
      fail("expected an integer");
      return 0;
    }
    return (long)v;
  }
  // one line (without its terminator and trailing CR)
  std::string line() {
    const char* s = p;
    while (p < end && *p != '\n') ++p;
    const char* e = p;
    if (p < end) ++p;
    while (e > s && e[-1] == '\r') --e;
    return std::string(s, e);
  }
  double fail(const char* what) {
    if (ok) g_io_error = file + ": " + what;
    ok = false;
    return 0.0;
  }
};

bool parse_line_number(Cursor& c, double& v, const char* what) {
  const std::string s = c.line();
  char* q = nullptr;
  v = std::strtod(s.c_str(), &q);
  if (q == s.c_str()) {
    g_io_error = c.file + ": line '" + s + "' is not a number (" + what + ")";
    return false;
  }
  return true;
}

}  // namespace

extern "C" ibem3d_inputs* ibem3d_read_inputs(const char* dir) {
  g_io_error.clear();
  if (!dir) {
    g_io_error = "ibem3d_read_inputs: dir is NULL";
    return nullptr;
  }
  const std::string d = std::string(dir) + "/";
  auto* in = new ibem3d_inputs();
  std::string txt;

  // position.txt
  if (!slurp(d + "position.txt", txt)) {
    g_io_error = "cannot open " + d + "position.txt";
    delete in;
    return nullptr;
  }
  {
    Cursor c{txt.data(), txt.data() + txt.size(), d + "position.txt"};
    const long num = c.integer();
    in->prob.E1 = c.num();
    in->prob.nu1 = c.num();
    if (c.ok && num < 0) c.fail("negative sphere count");
    if (c.ok) {
      in->centres.resize(3 * num);
      in->radii.resize(num);
      for (long k = 0; k < num && c.ok; ++k) {
        for (int j = 0; j < 3; ++j) in->centres[3 * k + j] = c.num();
        in->radii[k] = c.num();
      }
    }
    if (!c.ok) {
      delete in;
      return nullptr;
    }
    in->prob.num = (int)num;
  }

  // BEM.txt
  if (!slurp(d + "BEM.txt", txt)) {
    g_io_error = "cannot open " + d + "BEM.txt";
    delete in;
    return nullptr;
  }
  {
    Cursor c{txt.data(), txt.data() + txt.size(), d + "BEM.txt"};
    in->title = c.line();
    double E0, nu0, NNd, NEd;
    if (!parse_line_number(c, E0, "E0") || !parse_line_number(c, nu0, "nu0") ||
        !parse_line_number(c, NNd, "NN") || !parse_line_number(c, NEd, "NE")) {
      delete in;
      return nullptr;
    }
    const long NN = (long)NNd, NE = (long)NEd;
    if (NN <= 0 || NE <= 0) c.fail("node and element counts must be positive");
    if (c.ok) {
      in->nodes.resize(3 * NN);
      for (auto& v : in->nodes) v = c.num();
      in->conn.resize(4 * NE);
      for (auto& v : in->conn) v = (int)c.integer();
      in->bc.resize(6 * NN);
      for (auto& v : in->bc) v = c.num();
      in->tbc.resize(24 * NE);
      for (auto& v : in->tbc) v = c.num();
    }
    if (!c.ok) {
      delete in;
      return nullptr;
    }
    in->prob.NN = (int)NN;
    in->prob.NE = (int)NE;
    in->prob.E0 = E0;
    in->prob.nu0 = nu0;
  }

  // postprocess.txt (optional: no file means no stress points)
  if (slurp(d + "postprocess.txt", txt)) {
    Cursor c{txt.data(), txt.data() + txt.size(), d + "postprocess.txt"};
    const long nump = c.integer();
    if (c.ok && nump < 0) c.fail("negative point count");
    if (c.ok) {
      in->points.resize(3 * nump);
      for (auto& v : in->points) v = c.num();
    }
    if (!c.ok) {
      delete in;
      return nullptr;
    }
    in->prob.nump = (int)nump;
  }

  ibem3d_problem& p = in->prob;
  p.nodes = in->nodes.data();
  p.conn = in->conn.data();
  p.bc = in->bc.data();
  p.tbc = in->tbc.data();
  p.centres = in->centres.data();
  p.radii = in->radii.data();
  p.points = p.nump > 0 ? in->points.data() : nullptr;
  return in;
}

extern "C" const ibem3d_problem* ibem3d_inputs_problem(const ibem3d_inputs* in) {
  return in ? &in->prob : nullptr;
}

extern "C" const char* ibem3d_inputs_title(const ibem3d_inputs* in) {
  return in ? in->title.c_str() : "";
}

extern "C" void ibem3d_free_inputs(ibem3d_inputs* in) { delete in; }

extern "C" const char* ibem3d_io_error(void) { return g_io_error.c_str(); }

extern "C" int ibem3d_write_stress(const char* path, const double* s, int nump, int full) {
  if (!path || (nump > 0 && !s) || nump < 0) {
    g_io_error = "ibem3d_write_stress: bad argument";
    return IBEM3D_EINVAL;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    g_io_error = std::string("cannot write ") + path;
    return IBEM3D_EINVAL;
  }
  const char* fmt = full ? "%.17g" : "%.6e";
  for (int k = 0; k < nump; ++k) {
    for (int j = 0; j < 6; ++j) {
      if (j) std::fputc(' ', f);
      std::fprintf(f, fmt, s[6 * k + j]);
    }
    std::fputc('\n', f);
  }
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) {
    g_io_error = std::string("write error on ") + path;
    return IBEM3D_EINVAL;
  }
  return IBEM3D_OK;
}
