// graph_io.cpp -- .mtx reader with the reference's adjacency order.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "fap_dev.h"
#include "host.h"

namespace fapi {

std::string reference_graph_path(const std::string &name, bool directed, bool weighted) {
  if (name.size() > 4 && name.compare(name.size() - 4, 4, ".mtx") == 0) return name;
  std::string dir = std::string("graph/") + (weighted ? "weight-" : "unweight-") +
                    (directed ? "directed/" : "undirected/");
  return dir + name + ".mtx";
}

namespace {

struct Cursor {
  const char *p;
  const char *end;
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' ||
                       *p == '\f'))
      p++;
  }
  // istream >> int semantics for the inputs we accept: optional sign, digits.
  bool read_int(long long *v) {
    skip_ws();
    if (p >= end) return false;
    char *q = nullptr;
    errno = 0;
    long long x = strtoll(p, &q, 10);
    if (q == p || errno) return false;
    p = q;
    *v = x;
    return true;
  }
  bool read_double(double *v) {
    skip_ws();
    if (p >= end) return false;
    char *q = nullptr;
    double x = strtod(p, &q);
    if (q == p) return false;
    p = q;
    *v = x;
    return true;
  }
};

}  // namespace

int read_mtx(const std::string &path, bool directed, bool weighted, HostCSR *out,
             std::string *err) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    *err = "cannot open graph file '" + path + "'";
    return FAP_E_IO;
  }
  std::vector<char> buf;
  if (fseek(f, 0, SEEK_END) == 0) {
    long sz = ftell(f);
    if (sz > 0) {
      buf.resize((size_t)sz + 1);
      fseek(f, 0, SEEK_SET);
      size_t got = fread(buf.data(), 1, (size_t)sz, f);
      buf.resize(got);
    }
  }
  fclose(f);
  buf.push_back('\0');
  Cursor c{buf.data(), buf.data() + buf.size() - 1};
  // leading comment lines: only lines whose first character is '%'
  while (c.p < c.end && *c.p == '%') {
    const char *nl = (const char *)memchr(c.p, '\n', c.end - c.p);
    c.p = nl ? nl + 1 : c.end;
  }
  long long n = 0, t = 0, m = 0;
  if (!c.read_int(&n) || !c.read_int(&t) || !c.read_int(&m) || n < 0 || m < 0 ||
      n > INT32_MAX - 1) {
    *err = "malformed header in '" + path + "' (expected: n n m)";
    return FAP_E_IO;
  }
  std::vector<int32_t> au, av, aw;
  const size_t arcs_hint = (size_t)(directed ? m : 2 * m);
  au.reserve(arcs_hint);
  av.reserve(arcs_hint);
  aw.reserve(arcs_hint);
  for (;;) {
    long long a, b;
    const char *save = c.p;
    if (!c.read_int(&a)) break;
    if (!c.read_int(&b)) {
      c.p = save;
      break;
    }
    long long wv = 1;
    if (weighted) {
      double x;
      if (!c.read_double(&x)) break;
      x = fabs(x);
      if (x != floor(x) || x >= (double)FAP_INF) {
        *err = "weight " + std::to_string(x) + " is not an integer below INF (int32 build)";
        return FAP_E_IO;
      }
      wv = (long long)x;
    }
    if (a < 1 || a > n || b < 1 || b > n) {
      *err = "vertex id out of range 1.." + std::to_string(n) + " in '" + path + "'";
      return FAP_E_IO;
    }
    au.push_back((int32_t)(a - 1));
    av.push_back((int32_t)(b - 1));
    aw.push_back((int32_t)wv);
    if (!directed) {
      au.push_back((int32_t)(b - 1));
      av.push_back((int32_t)(a - 1));
      aw.push_back((int32_t)wv);
    }
  }
  if (au.size() > (size_t)INT32_MAX) {
    *err = "too many arcs for int32 CSR";
    return FAP_E_IO;
  }
  const int32_t V = (int32_t)n;
  const int64_t E = (int64_t)au.size();
  out->V = V;
  out->E = E;
  out->row_offset.assign((size_t)V + 1, 0);
  for (int64_t e = 0; e < E; e++) out->row_offset[au[e] + 1]++;
  for (int32_t v = 0; v < V; v++) out->row_offset[v + 1] += out->row_offset[v];
  out->col.resize((size_t)E);
  out->w.resize((size_t)E);
  std::vector<int32_t> pos(out->row_offset.begin(), out->row_offset.end() - 1);
  // head insertion => newest arc first: place arcs from the last one read
  for (int64_t e = E - 1; e >= 0; e--) {
    int32_t slot = pos[au[e]]++;
    out->col[slot] = av[e];
    out->w[slot] = aw[e];
  }
  return FAP_OK;
}

}  // namespace fapi
