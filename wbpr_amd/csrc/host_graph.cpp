// Host-side graph code of libmf: file loaders and writers, seeded instance generators,
// residual-layout builders and the preflow -> flow repair.  Pure C++ (no HIP).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <string>

#include "mf_internal.h"

namespace mf {

mf_host_graph *host_graph_alloc(int32_t n, int64_t m) {
  mf_host_graph *g = (mf_host_graph *)calloc(1, sizeof(mf_host_graph));
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  g->offsets = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
  g->dst = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  g->cap = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  g->s = g->t = -1;
  g->m_header = m;
  g->m_processed = m;
  if (!g->offsets || !g->dst || !g->cap) {
    free(g->offsets);
    free(g->dst);
    free(g->cap);
    free(g);
    return nullptr;
  }
  return g;
}

mf_status host_graph_from_arcs(int32_t n, const ArcList &arcs, int32_t s, int32_t t,
                               mf_host_graph **out) {
  const int64_t m = (int64_t)arcs.size();
  if (m > INT32_MAX) return MF_ERR_OVERFLOW;
  mf_host_graph *g = host_graph_alloc(n, m);
  if (!g) return MF_ERR_OUT_OF_MEMORY;
  std::vector<int32_t> cnt((size_t)n + 1, 0);
  for (int64_t i = 0; i < m; ++i) cnt[(size_t)arcs.u[i] + 1]++;
  g->offsets[0] = 0;
  for (int32_t v = 0; v < n; ++v) g->offsets[v + 1] = g->offsets[v] + cnt[(size_t)v + 1];
  std::vector<int32_t> pos(g->offsets, g->offsets + n);
  for (int64_t i = 0; i < m; ++i) {
    int32_t k = pos[(size_t)arcs.u[i]]++;
    g->dst[k] = arcs.v[i];
    g->cap[k] = arcs.c[i];
  }
  g->s = s;
  g->t = t;
  *out = g;
  return MF_OK;
}

mf_status validate_csr(const mf_csr *g) {
  if (!g || g->n < 2 || g->m < 0 || !g->offsets || (g->m > 0 && !g->dst))
    return MF_ERR_INVALID_ARGUMENT;
  if (g->m > INT32_MAX / 2) return MF_ERR_OVERFLOW;  // two residual slots per arc must fit int32
  if (g->s < 0 || g->s >= g->n || g->t < 0 || g->t >= g->n || g->s == g->t)
    return MF_ERR_INVALID_ARGUMENT;
  if (g->offsets[0] != 0 || (int64_t)g->offsets[g->n] != g->m) return MF_ERR_INVALID_ARGUMENT;
  for (int32_t v = 0; v < g->n; ++v)
    if (g->offsets[v + 1] < g->offsets[v]) return MF_ERR_INVALID_ARGUMENT;
  for (int64_t i = 0; i < g->m; ++i) {
    if (g->dst[i] < 0 || g->dst[i] >= g->n) return MF_ERR_INVALID_ARGUMENT;
    if (g->cap && g->cap[i] < 0) return MF_ERR_INVALID_ARGUMENT;
  }
  return MF_OK;
}

// ---------------------------------------------------------------------------------------------
// File reading helpers
// ---------------------------------------------------------------------------------------------
static bool read_whole_file(const char *path, std::string &buf) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return false;
  }
  long sz = ftell(f);
  if (sz < 0) {
    fclose(f);
    return false;
  }
  rewind(f);
  buf.resize((size_t)sz);
  size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, f) : 0;
  fclose(f);
  return got == (size_t)sz;
}

// Parses a signed decimal integer at p (skipping blanks); returns false if none.
static inline bool parse_int(const char *&p, const char *end, int64_t &out) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) {
    neg = (*p == '-');
    ++p;
  }
  if (p >= end || !isdigit((unsigned char)*p)) return false;
  int64_t v = 0;
  while (p < end && isdigit((unsigned char)*p)) {
    v = v * 10 + (*p - '0');
    if (v > (int64_t)1 << 40) return false;
    ++p;
  }
  out = neg ? -v : v;
  return true;
}

static inline bool parse_word(const char *&p, const char *end, std::string &w) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  const char *b = p;
  while (p < end && !isspace((unsigned char)*p)) ++p;
  w.assign(b, p);
  return p > b;
}

}  // namespace mf

using namespace mf;

extern "C" {

void mf_host_graph_free(mf_host_graph *g) {
  if (!g) return;
  free(g->offsets);
  free(g->dst);
  free(g->cap);
  free(g);
}

mf_csr mf_host_graph_view(const mf_host_graph *g) {
  mf_csr c;
  memset(&c, 0, sizeof(c));
  if (!g) return c;
  c.n = g->n;
  c.m = g->m;
  c.offsets = g->offsets;
  c.dst = g->dst;
  c.cap = g->cap;
  c.s = g->s;
  c.t = g->t;
  return c;
}

// SNAP text edge list (CSRGraph::buildFromTxtFile, maxflow-cuda/src/graph.cpp:28-67):
// n and m come from the '# Nodes: N Edges: M' header, every other '#' line is a comment,
// each remaining line is "from to"; adjacency of every vertex is sorted; capacities are 1.
mf_status mf_load_snap_txt(const char *path, mf_host_graph **out) {
  if (!path || !out) return MF_ERR_INVALID_ARGUMENT;
  std::string buf;
  if (!read_whole_file(path, buf)) return MF_ERR_IO;
  const char *p = buf.data(), *end = p + buf.size();
  int64_t n = -1, m_header = -1, cnt = 0;
  std::vector<int64_t> from, to;
  while (p < end) {
    const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
    if (!eol) eol = end;
    const char *q = p;
    if (q < eol && *q == '#') {
      const std::string line(q, eol);
      const size_t at = line.find("# Nodes:");
      long long nn = 0, mm = 0;
      if (at != std::string::npos &&
          sscanf(line.c_str() + at, "# Nodes: %lld Edges: %lld", &nn, &mm) == 2) {
        n = nn;
        m_header = mm;
      }
    } else {
      const char *r = q;
      while (r < eol && isspace((unsigned char)*r)) ++r;
      if (r < eol) {
        int64_t a, b;
        if (!parse_int(q, eol, a) || !parse_int(q, eol, b)) return MF_ERR_PARSE;
        from.push_back(a);
        to.push_back(b);
        ++cnt;
      }
    }
    p = eol + 1;
  }
  if (n < 2 || n > INT32_MAX - 1) return MF_ERR_PARSE;
  ArcList arcs;
  arcs.u.reserve(from.size());
  // Per-vertex sorted adjacency (graph.cpp:57): sort arcs by (from, to).
  std::vector<size_t> idx(from.size());
  std::iota(idx.begin(), idx.end(), 0);
  for (size_t i = 0; i < from.size(); ++i)
    if (from[i] < 0 || from[i] >= n || to[i] < 0 || to[i] >= n) return MF_ERR_PARSE;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
    return from[x] != from[y] ? from[x] < from[y] : to[x] < to[y];
  });
  for (size_t k : idx) arcs.add((int32_t)from[k], (int32_t)to[k], 1);
  mf_status st = host_graph_from_arcs((int32_t)n, arcs, -1, -1, out);
  if (st != MF_OK) return st;
  (*out)->m_header = m_header;
  (*out)->m_processed = cnt;
  return MF_OK;
}

// SNAP binary (CSRGraph::loadFromBinary, maxflow-cuda/src/graph.cpp:292-314):
// int32 [n][m][m_processed][dst x m_processed][offsets x (n+1)], unit capacities.
mf_status mf_load_snap_bin(const char *path, mf_host_graph **out) {
  if (!path || !out) return MF_ERR_INVALID_ARGUMENT;
  FILE *f = fopen(path, "rb");
  if (!f) return MF_ERR_IO;
  int32_t hdr[3];
  if (fread(hdr, sizeof(int32_t), 3, f) != 3) {
    fclose(f);
    return MF_ERR_PARSE;
  }
  int32_t n = hdr[0], m_hdr = hdr[1], mp = hdr[2];
  if (n < 2 || mp < 0 || n == INT32_MAX) {
    fclose(f);
    return MF_ERR_PARSE;
  }
  // the arrays the header announces must be in the file: a truncated or foreign file fails here,
  // before its header sizes the allocation
  const long at = ftell(f);
  if (at < 0 || fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return MF_ERR_IO;
  }
  const long size = ftell(f);
  if (size < at + (long)sizeof(int32_t) * ((long)mp + (long)n + 1) || fseek(f, at, SEEK_SET) != 0) {
    fclose(f);
    return MF_ERR_PARSE;
  }
  mf_host_graph *g = host_graph_alloc(n, mp);
  if (!g) {
    fclose(f);
    return MF_ERR_OUT_OF_MEMORY;
  }
  bool ok = fread(g->dst, sizeof(int32_t), (size_t)mp, f) == (size_t)mp &&
            fread(g->offsets, sizeof(int32_t), (size_t)n + 1, f) == (size_t)n + 1;
  fclose(f);
  if (ok) {
    ok = g->offsets[0] == 0 && g->offsets[n] == mp;
    for (int32_t v = 0; ok && v < n; ++v) ok = g->offsets[v] <= g->offsets[v + 1];
    for (int32_t i = 0; ok && i < mp; ++i) ok = g->dst[i] >= 0 && g->dst[i] < n;
  }
  if (!ok) {
    mf_host_graph_free(g);
    return MF_ERR_PARSE;
  }
  for (int32_t i = 0; i < mp; ++i) g->cap[i] = 1;
  g->m_header = m_hdr;
  g->m_processed = mp;
  *out = g;
  return MF_OK;
}

// CSRGraph::saveToBinary (maxflow-cuda/src/graph.cpp:219-233).
mf_status mf_save_snap_bin(const char *path, const mf_csr *g) {
  if (!path || !g || !g->offsets || g->n < 1 || g->m < 0 || g->m > INT32_MAX)
    return MF_ERR_INVALID_ARGUMENT;
  FILE *f = fopen(path, "wb");
  if (!f) return MF_ERR_IO;
  int32_t hdr[3] = {g->n, (int32_t)g->m, (int32_t)g->m};
  bool ok = fwrite(hdr, sizeof(int32_t), 3, f) == 3 &&
            fwrite(g->dst, sizeof(int32_t), (size_t)g->m, f) == (size_t)g->m &&
            fwrite(g->offsets, sizeof(int32_t), (size_t)g->n + 1, f) == (size_t)g->n + 1;
  ok = (fclose(f) == 0) && ok;
  return ok ? MF_OK : MF_ERR_IO;
}

// DIMACS max-flow format (CSRGraph::buildFromDIMACSFile, maxflow-cuda/src/graph.cpp:151-214):
// 'c' comments, 'p max n m', 'n id s|t', 'a u v cap', 1-based ids; CSR in vertex order with
// arcs of a vertex in file order.
mf_status mf_load_dimacs(const char *path, mf_host_graph **out) {
  if (!path || !out) return MF_ERR_INVALID_ARGUMENT;
  std::string buf;
  if (!read_whole_file(path, buf)) return MF_ERR_IO;
  const char *p = buf.data(), *end = p + buf.size();
  int64_t n = -1, m_header = -1, s = -1, t = -1;
  ArcList raw;
  std::string word;
  while (p < end) {
    const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
    if (!eol) eol = end;
    const char *q = p;
    if (q < eol) {
      char kind = *q;
      if (kind == 'p') {
        ++q;
        if (!parse_word(q, eol, word)) return MF_ERR_PARSE;
        if (!parse_int(q, eol, n) || !parse_int(q, eol, m_header)) return MF_ERR_PARSE;
        if (n < 2 || n > INT32_MAX - 1 || m_header < 0) return MF_ERR_PARSE;
        raw.u.reserve((size_t)std::min<int64_t>(m_header, (int64_t)1 << 31));
        raw.v.reserve(raw.u.capacity());
        raw.c.reserve(raw.u.capacity());
      } else if (kind == 'n') {
        ++q;
        int64_t id;
        if (!parse_int(q, eol, id) || !parse_word(q, eol, word)) return MF_ERR_PARSE;
        if (word == "s") s = id - 1;
        else if (word == "t") t = id - 1;
      } else if (kind == 'a') {
        ++q;
        int64_t u, v, c;
        if (!parse_int(q, eol, u) || !parse_int(q, eol, v) || !parse_int(q, eol, c))
          return MF_ERR_PARSE;
        if (c < 0 || c > INT32_MAX) return MF_ERR_PARSE;
        if (u < 1 || v < 1 || u > INT32_MAX || v > INT32_MAX) return MF_ERR_PARSE;
        raw.add((int32_t)(u - 1), (int32_t)(v - 1), (int32_t)c);
      }
    }
    p = eol + 1;
  }
  if (n < 0) return MF_ERR_PARSE;
  for (size_t i = 0; i < raw.size(); ++i)
    if (raw.u[i] >= n || raw.v[i] >= n) return MF_ERR_PARSE;
  if (s >= n || t >= n) return MF_ERR_PARSE;
  mf_status st = host_graph_from_arcs((int32_t)n, raw, (int32_t)s, (int32_t)t, out);
  if (st != MF_OK) return st;
  (*out)->m_header = m_header;
  (*out)->m_processed = (int64_t)raw.size();
  return MF_OK;
}

mf_status mf_save_dimacs(const char *path, const mf_csr *g) {
  if (!path || !g || !g->offsets || g->n < 1 || g->m < 0) return MF_ERR_INVALID_ARGUMENT;
  FILE *f = fopen(path, "w");
  if (!f) return MF_ERR_IO;
  fprintf(f, "p max %d %lld\n", g->n, (long long)g->m);
  if (g->s >= 0) fprintf(f, "n %d s\n", g->s + 1);
  if (g->t >= 0) fprintf(f, "n %d t\n", g->t + 1);
  for (int32_t u = 0; u < g->n; ++u)
    for (int32_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i)
      fprintf(f, "a %d %d %d\n", u + 1, g->dst[i] + 1, g->cap ? g->cap[i] : 1);
  return fclose(f) == 0 ? MF_OK : MF_ERR_IO;
}

// ---------------------------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------------------------

// Washington random level graph, RLevel(d1, d2, r) (network/DIMACS/washington/washington.c:243-279):
// vertex 0 is the source, level i (0-based) vertex j (1..d1) is i*d1 + j, the sink is d1*d2+1.
// Source -> level 0 and last level -> sink arcs have capacity 3r; every other vertex has three
// distinct uniform targets in the next level (RandomSubset) with capacity uniform [1, r].
mf_status mf_gen_washington_rlg(int32_t d1, int32_t d2, int32_t r, uint64_t seed,
                                mf_host_graph **out) {
  if (!out || d1 < 3 || d2 < 2 || r < 1 || (int64_t)d1 * d2 + 2 > INT32_MAX ||
      (int64_t)3 * r > INT32_MAX)
    return MF_ERR_INVALID_ARGUMENT;
  Rng rng(seed);
  const int32_t n = d1 * d2 + 2, source = 0, sink = d1 * d2 + 1;
  ArcList arcs;
  arcs.u.reserve((size_t)2 * d1 + (size_t)3 * d1 * (d2 - 1));
  for (int32_t i = 1; i <= d1; ++i) {
    arcs.add(source, source + i, 3 * r);
    arcs.add(sink - i, sink, 3 * r);
  }
  for (int32_t i = 0; i < d2 - 1; ++i) {
    for (int32_t j = 1; j <= d1; ++j) {
      int32_t x[3];
      for (int k = 0; k < 3; ++k) {  // three distinct values in [1, d1]
        bool again;
        do {
          x[k] = (int32_t)rng.range(1, d1);
          again = false;
          for (int q = 0; q < k; ++q) again |= (x[q] == x[k]);
        } while (again);
      }
      for (int k = 0; k < 3; ++k)
        arcs.add(i * d1 + j, (i + 1) * d1 + x[k], (int32_t)rng.range(1, r));
    }
  }
  return host_graph_from_arcs(n, arcs, source, sink, out);
}

// Genrmf (network/DIMACS/genrmf/genrmf.c:76-161): b frames of a x a grids; in-frame
// 4-neighbour arcs both ways with capacity c2*a*a; frame z -> z+1 arcs through a random
// permutation (Fisher-Yates applied cumulatively, genrmf.c:140-150) with capacity uniform
// [c1, c2]; source = first vertex of frame 0, sink = last vertex of frame b-1.
mf_status mf_gen_genrmf(int32_t a, int32_t b, int32_t c1, int32_t c2, uint64_t seed,
                        mf_host_graph **out) {
  if (!out || a < 2 || b < 1 || c1 < 0 || c2 < c1) return MF_ERR_INVALID_ARGUMENT;
  const int64_t AA = (int64_t)a * a, n64 = AA * b;
  if (n64 > INT32_MAX || (int64_t)c2 * AA > INT32_MAX) return MF_ERR_OVERFLOW;
  const int32_t C2AA = (int32_t)(c2 * AA);
  Rng rng(seed);
  std::vector<int32_t> perm((size_t)AA + 1);
  for (int64_t x = 1; x <= AA; ++x) perm[(size_t)x] = (int32_t)x;
  ArcList arcs;
  arcs.u.reserve((size_t)(5 * AA * b - 4 * (int64_t)a * b - AA));
  for (int32_t z = 1; z <= b; ++z) {
    const int64_t offset = AA * (z - 1);
    if (z != b) {
      for (int64_t i = 1; i < AA; ++i) {
        int64_t j = rng.range(i, AA);
        std::swap(perm[(size_t)i], perm[(size_t)j]);
      }
    }
    for (int32_t x = 1; x <= a; ++x) {
      for (int32_t y = 1; y <= a; ++y) {
        const int64_t cv = offset + (int64_t)(x - 1) * a + y;  // 1-based
        auto conn = [&](int32_t x1, int32_t y1) {
          int64_t cv1 = offset + (int64_t)(x1 - 1) * a + y1;
          arcs.add((int32_t)(cv - 1), (int32_t)(cv1 - 1), C2AA);
        };
        if (z != b)
          arcs.add((int32_t)(cv - 1), (int32_t)(offset + AA + perm[(size_t)(cv - offset)] - 1),
                   (int32_t)rng.range(c1, c2));
        if (y < a) conn(x, y + 1);
        if (y > 1) conn(x, y - 1);
        if (x < a) conn(x + 1, y);
        if (x > 1) conn(x - 1, y);
      }
    }
  }
  return host_graph_from_arcs((int32_t)n64, arcs, 0, (int32_t)(n64 - 1), out);
}

// R-MAT (GTgraph/R-MAT/graph.c:158-182 quadrant recursion, without the per-level parameter
// noise of graph.c:184-217): edge_factor * 2^scale samples; self-loops and duplicates are
// dropped; capacities uniform [1, cap_max]; s = max out-degree vertex, t = max in-degree vertex
// (lowest id on ties; if equal to s, the next best in-degree).
mf_status mf_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      int32_t cap_max, uint64_t seed, mf_host_graph **out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || cap_max < 1 || a < 0 || b < 0 ||
      c < 0 || a + b + c > 1.0)
    return MF_ERR_INVALID_ARGUMENT;
  const int64_t n = (int64_t)1 << scale;
  const int64_t samples = n * edge_factor;
  if (n > INT32_MAX || samples > INT32_MAX) return MF_ERR_OVERFLOW;
  Rng rng(seed);
  std::vector<uint64_t> keys;
  keys.reserve((size_t)samples);
  const double ab = a + b, abc = a + b + c;
  for (int64_t e = 0; e < samples; ++e) {
    uint64_t u = 0, v = 0;
    for (int32_t lvl = scale - 1; lvl >= 0; --lvl) {
      double p = rng.uniform();
      if (p < a) {
      } else if (p < ab) {
        v |= (1ull << lvl);
      } else if (p < abc) {
        u |= (1ull << lvl);
      } else {
        u |= (1ull << lvl);
        v |= (1ull << lvl);
      }
    }
    if (u != v) keys.push_back((u << 32) | v);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  mf_host_graph *g = host_graph_alloc((int32_t)n, (int64_t)keys.size());
  if (!g) return MF_ERR_OUT_OF_MEMORY;
  std::vector<int32_t> indeg((size_t)n, 0);
  std::fill(g->offsets, g->offsets + n + 1, 0);
  for (size_t i = 0; i < keys.size(); ++i) {
    int32_t u = (int32_t)(keys[i] >> 32), v = (int32_t)(keys[i] & 0xffffffffu);
    g->offsets[u + 1]++;
    g->dst[i] = v;
    g->cap[i] = (int32_t)rng.range(1, cap_max);
    indeg[(size_t)v]++;
  }
  for (int64_t v = 0; v < n; ++v) g->offsets[v + 1] += g->offsets[v];
  int32_t s = 0, t = -1;
  for (int32_t v = 1; v < n; ++v)
    if (g->offsets[v + 1] - g->offsets[v] > g->offsets[s + 1] - g->offsets[s]) s = v;
  for (int32_t v = 0; v < n; ++v) {
    if (v == s) continue;
    if (t < 0 || indeg[(size_t)v] > indeg[(size_t)t]) t = v;
  }
  g->s = s;
  g->t = t;
  *out = g;
  return MF_OK;
}

// Bipartite matching network: source 0 -> every left vertex (1..L), each left vertex ->
// Poisson(mean_deg) distinct uniform right vertices (L+1..L+R), every right vertex -> sink
// (L+R+1); all capacities 1.
mf_status mf_gen_bipartite(int32_t L, int32_t R, double mean_deg, uint64_t seed,
                           mf_host_graph **out) {
  if (!out || L < 1 || R < 1 || mean_deg < 0 || (int64_t)L + R + 2 > INT32_MAX)
    return MF_ERR_INVALID_ARGUMENT;
  Rng rng(seed);
  const int32_t n = L + R + 2, source = 0, sink = L + R + 1;
  ArcList arcs;
  arcs.u.reserve((size_t)L + R + (size_t)(L * (mean_deg + 1)));
  const double lim = std::exp(-mean_deg);
  std::vector<int32_t> pick;
  for (int32_t l = 1; l <= L; ++l) arcs.add(source, l, 1);
  for (int32_t l = 1; l <= L; ++l) {
    // Poisson sample by Knuth's product method
    int32_t k = 0;
    double prod = rng.uniform();
    while (prod > lim) {
      ++k;
      prod *= rng.uniform();
    }
    k = std::min(k, R);
    pick.clear();
    while ((int32_t)pick.size() < k) {
      int32_t r = (int32_t)rng.range(1, R);
      if (std::find(pick.begin(), pick.end(), r) == pick.end()) pick.push_back(r);
    }
    for (int32_t r : pick) arcs.add(l, L + r, 1);
  }
  for (int32_t r = 1; r <= R; ++r) arcs.add(L + r, sink, 1);
  return host_graph_from_arcs(n, arcs, source, sink, out);
}

// KONECT bipartite edge list (the paper's bipartite-matching inputs, p.4): '%' comment lines,
// then "u v [weight [time]]" with 1-based ids in two separate id spaces (left u, right v).
// Matching network: super-source 0 -> every left vertex 1..L, left u -> right L+v for every
// edge, every right vertex -> super-sink L+R+1; all capacities 1.  L and R are the largest ids
// seen (or the counts of a "% m L R" size line, when larger).
mf_status mf_load_konect_bipartite(const char *path, mf_host_graph **out) {
  if (!path || !out) return MF_ERR_INVALID_ARGUMENT;
  std::string buf;
  if (!read_whole_file(path, buf)) return MF_ERR_IO;
  const char *p = buf.data(), *end = p + buf.size();
  std::vector<int32_t> eu, ev;
  int64_t L = 0, R = 0;
  int comment_lines = 0;
  while (p < end) {
    const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
    if (!eol) eol = end;
    const char *q = p;
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q < eol && *q == '%') {
      // the second comment line of a KONECT file holds "m L R"
      if (++comment_lines == 2) {
        ++q;
        int64_t a, b, c;
        if (parse_int(q, eol, a) && parse_int(q, eol, b) && parse_int(q, eol, c)) {
          L = std::max(L, b);
          R = std::max(R, c);
        }
      }
    } else if (q < eol) {
      int64_t u, v;
      if (!parse_int(q, eol, u) || !parse_int(q, eol, v)) return MF_ERR_PARSE;
      if (u < 1 || v < 1 || u > INT32_MAX || v > INT32_MAX) return MF_ERR_PARSE;
      eu.push_back((int32_t)u);
      ev.push_back((int32_t)v);
      L = std::max(L, u);
      R = std::max(R, v);
    }
    p = eol + 1;
  }
  if (L + R + 2 > INT32_MAX) return MF_ERR_OVERFLOW;
  const int32_t n = (int32_t)(L + R + 2), source = 0, sink = n - 1;
  ArcList arcs;
  for (int32_t l = 1; l <= L; ++l) arcs.add(source, l, 1);
  for (size_t i = 0; i < eu.size(); ++i) arcs.add(eu[i], (int32_t)L + ev[i], 1);
  for (int32_t r = 1; r <= R; ++r) arcs.add((int32_t)L + r, sink, 1);
  mf_status st = host_graph_from_arcs(n, arcs, source, sink, out);
  if (st != MF_OK) return st;
  (*out)->m_header = (int64_t)eu.size();
  (*out)->m_processed = (int64_t)eu.size();
  return MF_OK;
}

// Super-source n and super-sink n+1 joined to the given sources and sinks (the paper's SNAP
// transform, p.4).  The arc super-source -> s_i carries the total capacity out of s_i and the
// arc t_i -> super-sink the total capacity into t_i, so neither limits the flow.
mf_status mf_super_terminals(const mf_csr *g, const int32_t *sources, int32_t ns,
                             const int32_t *sinks, int32_t nt, mf_host_graph **out) {
  if (!g || !out || !sources || !sinks || ns < 1 || nt < 1) return MF_ERR_INVALID_ARGUMENT;
  if (!g->offsets || !g->dst || g->n < 1 || g->m < 0) return MF_ERR_INVALID_ARGUMENT;
  if ((int64_t)g->n + 2 > INT32_MAX) return MF_ERR_OVERFLOW;
  std::vector<char> role((size_t)g->n, 0);
  for (int32_t i = 0; i < ns; ++i) {
    if (sources[i] < 0 || sources[i] >= g->n || role[(size_t)sources[i]]) return MF_ERR_INVALID_ARGUMENT;
    role[(size_t)sources[i]] = 1;
  }
  for (int32_t i = 0; i < nt; ++i) {
    if (sinks[i] < 0 || sinks[i] >= g->n || role[(size_t)sinks[i]]) return MF_ERR_INVALID_ARGUMENT;
    role[(size_t)sinks[i]] = 2;
  }
  std::vector<int64_t> cin((size_t)g->n, 0), cout((size_t)g->n, 0);
  ArcList arcs;
  arcs.u.reserve((size_t)g->m + ns + nt);
  for (int32_t u = 0; u < g->n; ++u)
    for (int32_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) {
      const int32_t c = g->cap ? g->cap[a] : 1;
      arcs.add(u, g->dst[a], c);
      cout[(size_t)u] += c;
      cin[(size_t)g->dst[a]] += c;
    }
  const int32_t S = g->n, T = g->n + 1;
  int64_t total = 0;
  for (int32_t i = 0; i < ns; ++i) {
    const int64_t c = cout[(size_t)sources[i]];
    if (c > INT32_MAX) return MF_ERR_OVERFLOW;
    total += c;
    arcs.add(S, sources[i], (int32_t)c);
  }
  if (total > INT32_MAX) return MF_ERR_OVERFLOW;  // the super-source's excess must fit int32
  for (int32_t i = 0; i < nt; ++i) {
    const int64_t c = cin[(size_t)sinks[i]];
    if (c > INT32_MAX) return MF_ERR_OVERFLOW;
    arcs.add(sinks[i], T, (int32_t)c);
  }
  return host_graph_from_arcs(g->n + 2, arcs, S, T, out);
}

// Picks `pairs` source/sink pairs the way the paper describes for SNAP graphs (p.4): a BFS from
// a random vertex, the sink drawn from the vertices at the longest quarter of the BFS distances
// it reached; every vertex is used at most once.  Then joins them with mf_super_terminals.
mf_status mf_bfs_terminal_pairs(const mf_csr *g, int32_t pairs, uint64_t seed,
                                int32_t *sources_out, int32_t *sinks_out, mf_host_graph **out) {
  if (!g || !out || pairs < 1 || !g->offsets || !g->dst || g->n < 2) return MF_ERR_INVALID_ARGUMENT;
  const int32_t n = g->n;
  Rng rng(seed);
  std::vector<int32_t> dist((size_t)n), queue, srcs, snks, far;
  std::vector<char> used((size_t)n, 0);
  const int64_t tries = (int64_t)pairs * 64 + 256;
  for (int64_t attempt = 0; attempt < tries && (int32_t)srcs.size() < pairs; ++attempt) {
    const int32_t r = (int32_t)rng.range(0, n - 1);
    if (used[(size_t)r] || g->offsets[r + 1] == g->offsets[r]) continue;
    std::fill(dist.begin(), dist.end(), -1);
    queue.assign(1, r);
    dist[(size_t)r] = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int32_t u = queue[h];
      for (int32_t a = g->offsets[u]; a < g->offsets[u + 1]; ++a) {
        const int32_t w = g->dst[a];
        if (dist[(size_t)w] < 0 && (!g->cap || g->cap[a] > 0)) {
          dist[(size_t)w] = dist[(size_t)u] + 1;
          queue.push_back(w);
        }
      }
    }
    if (queue.size() < 2) continue;
    // distances of the reached vertices are non-decreasing along the queue: the last quarter of
    // the distance range holds the farthest vertices
    const int32_t dmax = dist[(size_t)queue.back()];
    const int32_t dcut = dmax - dmax / 4;
    far.clear();
    for (size_t h = queue.size(); h-- > 1;) {
      if (dist[(size_t)queue[h]] < std::max(dcut, 1)) break;
      if (!used[(size_t)queue[h]]) far.push_back(queue[h]);
    }
    if (far.empty()) continue;
    const int32_t t = far[(size_t)rng.range(0, (int64_t)far.size() - 1)];
    used[(size_t)r] = used[(size_t)t] = 1;
    srcs.push_back(r);
    snks.push_back(t);
  }
  if ((int32_t)srcs.size() < pairs) return MF_ERR_NO_PATH;
  if (sources_out) std::copy(srcs.begin(), srcs.end(), sources_out);
  if (sinks_out) std::copy(snks.begin(), snks.end(), sinks_out);
  return mf_super_terminals(g, srcs.data(), pairs, snks.data(), pairs, out);
}

}  // extern "C"

namespace mf {

// BCSR / BCSR2 slot structure.  Arcs are visited in CSR order; arc i = (u -> v, c) appends
// its forward slot (dst v, residual c) to u and its reverse slot (dst u, residual 0) to v.
// That is the append order of the BCSR DIMACS loader (maxflow-bcsr/src/graph.cpp:171,197-200)
// for a file whose arcs are grouped by tail, and the reverse slot of every arc is known
// when it is placed, so no backward-edge search (maxflow-bcsr/src/graph.cpp:351-362) is needed.
mf_status build_slot_layout(const mf_csr *g, SlotLayout &L) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  if (2 * m > INT32_MAX) return MF_ERR_OVERFLOW;
  L.n = n;
  L.slots = 2 * m;
  L.off.assign((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; ++u) {
    for (int32_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i) {
      L.off[(size_t)u + 1]++;
      L.off[(size_t)g->dst[i] + 1]++;
    }
  }
  for (int32_t v = 0; v < n; ++v) L.off[(size_t)v + 1] += L.off[(size_t)v];
  L.dst.resize((size_t)L.slots);
  L.rev.resize((size_t)L.slots);
  L.cap0.resize((size_t)L.slots);
  L.fslot.resize((size_t)m);
  std::vector<int32_t> pos(L.off.begin(), L.off.end() - 1);
  for (int32_t u = 0; u < n; ++u) {
    for (int32_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i) {
      const int32_t v = g->dst[i];
      const int32_t a = pos[(size_t)u]++;
      const int32_t b = pos[(size_t)v]++;
      L.dst[(size_t)a] = v;
      L.dst[(size_t)b] = u;
      L.rev[(size_t)a] = b;
      L.rev[(size_t)b] = a;
      L.cap0[(size_t)a] = g->cap ? g->cap[i] : 1;
      L.cap0[(size_t)b] = 0;
      L.fslot[(size_t)i] = a;
    }
  }
  return MF_OK;
}

// RCSR reverse CSR (maxflow-cuda/src/graph.cpp:352-398): in-arcs of every vertex grouped in
// CSR order; flow_index[k] is the arc itself.  The reference searches out(v) for the first arc
// to u, which is the same arc whenever the graph has no parallel arcs.
mf_status build_rcsr_layout(const mf_csr *g, RcsrLayout &L) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  L.n = n;
  L.m = m;
  L.roff.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < m; ++i) L.roff[(size_t)g->dst[i] + 1]++;
  for (int32_t v = 0; v < n; ++v) L.roff[(size_t)v + 1] += L.roff[(size_t)v];
  L.rdst.resize((size_t)m);
  L.fidx.resize((size_t)m);
  std::vector<int32_t> pos(L.roff.begin(), L.roff.end() - 1);
  for (int32_t u = 0; u < n; ++u) {
    for (int32_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i) {
      const int32_t k = pos[(size_t)g->dst[i]]++;
      L.rdst[(size_t)k] = u;
      L.fidx[(size_t)k] = i;
    }
  }
  return MF_OK;
}

// Preflow -> flow.  Cancels flow cycles with an iterative DFS over arcs carrying flow, then
// walks vertices children-first (DFS post-order of the acyclic flow graph) and removes each
// vertex's excess from its incoming flow, which moves the excess one step towards the source.
void return_excess_to_source(const mf_csr *g, int32_t *flow) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  std::vector<int32_t> tail((size_t)m);
  for (int32_t u = 0; u < n; ++u)
    for (int32_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i) {
      tail[(size_t)i] = u;
      if (g->dst[i] == u) flow[i] = 0;  // self-loops carry nothing useful
    }
  // 1. cycle cancelling
  std::vector<uint8_t> state((size_t)n, 0);  // 0 new, 1 on stack, 2 finished
  std::vector<int32_t> it((size_t)n), parent_arc((size_t)n, -1), stack;
  std::vector<int32_t> post;
  post.reserve((size_t)n);
  for (int32_t v = 0; v < n; ++v) it[(size_t)v] = g->offsets[v];
  for (int32_t root = 0; root < n; ++root) {
    if (state[(size_t)root] != 0) continue;
    stack.push_back(root);
    state[(size_t)root] = 1;
    while (!stack.empty()) {
      const int32_t v = stack.back();
      bool descended = false;
      while (it[(size_t)v] < g->offsets[v + 1]) {
        const int32_t i = it[(size_t)v];
        const int32_t w = g->dst[i];
        if (flow[i] <= 0 || state[(size_t)w] == 2) {
          ++it[(size_t)v];
          continue;
        }
        if (state[(size_t)w] == 0) {
          parent_arc[(size_t)w] = i;
          state[(size_t)w] = 1;
          stack.push_back(w);
          descended = true;
          break;
        }
        // w is on the stack: cycle w -> ... -> v -> w
        int32_t delta = flow[i];
        for (int32_t x = v; x != w; x = tail[(size_t)parent_arc[(size_t)x]])
          delta = std::min(delta, flow[parent_arc[(size_t)x]]);
        flow[i] -= delta;
        for (int32_t x = v; x != w; x = tail[(size_t)parent_arc[(size_t)x]])
          flow[parent_arc[(size_t)x]] -= delta;
        // unwind to the tail of the first saturated (now empty) arc on the cycle from w
        int32_t cut_tail = -1;
        {
          // walk the stack from w upwards
          size_t pos_w = stack.size() - 1;
          while (stack[pos_w] != w) --pos_w;
          for (size_t k = pos_w + 1; k < stack.size(); ++k) {
            if (flow[parent_arc[(size_t)stack[k]]] == 0) {
              cut_tail = stack[k - 1];
              break;
            }
          }
          if (cut_tail < 0) cut_tail = v;  // only the closing arc emptied
          while (stack.back() != cut_tail) {
            state[(size_t)stack.back()] = 0;
            stack.pop_back();
          }
        }
        descended = true;  // re-examine the new top
        break;
      }
      if (descended) continue;
      state[(size_t)v] = 2;
      post.push_back(v);
      stack.pop_back();
    }
  }
  // 2. excess removal, children first
  std::vector<int64_t> excess((size_t)n, 0);
  for (int64_t i = 0; i < m; ++i) {
    excess[(size_t)g->dst[i]] += flow[i];
    excess[(size_t)tail[(size_t)i]] -= flow[i];
  }
  std::vector<int32_t> roff((size_t)n + 1, 0), rarc((size_t)m);
  for (int64_t i = 0; i < m; ++i) roff[(size_t)g->dst[i] + 1]++;
  for (int32_t v = 0; v < n; ++v) roff[(size_t)v + 1] += roff[(size_t)v];
  {
    std::vector<int32_t> pos(roff.begin(), roff.end() - 1);
    for (int64_t i = 0; i < m; ++i) rarc[(size_t)pos[(size_t)g->dst[i]]++] = (int32_t)i;
  }
  for (int32_t v : post) {
    if (v == g->s || v == g->t) continue;
    for (int32_t k = roff[(size_t)v]; k < roff[(size_t)v + 1] && excess[(size_t)v] > 0; ++k) {
      const int32_t i = rarc[(size_t)k];
      if (flow[i] <= 0) continue;
      const int32_t d = (int32_t)std::min<int64_t>(flow[i], excess[(size_t)v]);
      flow[i] -= d;
      excess[(size_t)v] -= d;
      excess[(size_t)tail[(size_t)i]] += d;
    }
  }
}

}  // namespace mf

extern "C" {

mf_status mf_build_layout(const mf_csr *g, mf_layout layout, mf_layout_arrays *out) {
  if (!out) return MF_ERR_INVALID_ARGUMENT;
  memset(out, 0, sizeof(*out));
  mf_status st = validate_csr(g);
  if (st != MF_OK) return st;
  auto dup = [](const std::vector<int32_t> &v) {
    int32_t *p = (int32_t *)malloc(sizeof(int32_t) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(int32_t) * v.size());
    return p;
  };
  out->n = g->n;
  if (layout == MF_LAYOUT_RCSR) {
    RcsrLayout L;
    st = build_rcsr_layout(g, L);
    if (st != MF_OK) return st;
    out->slots = L.m;
    out->a = dup(L.roff);
    out->b = dup(L.rdst);
    out->c = dup(L.fidx);
  } else if (layout == MF_LAYOUT_BCSR || layout == MF_LAYOUT_BCSR2) {
    SlotLayout L;
    st = build_slot_layout(g, L);
    if (st != MF_OK) return st;
    out->slots = L.slots;
    out->a = dup(L.off);
    out->b = dup(L.dst);
    out->c = dup(L.rev);
    out->d = dup(L.cap0);
    out->e = dup(L.fslot);
  } else {
    return MF_ERR_INVALID_ARGUMENT;
  }
  return MF_OK;
}

void mf_layout_free(mf_layout_arrays *arr) {
  if (!arr) return;
  free(arr->a);
  free(arr->b);
  free(arr->c);
  free(arr->d);
  free(arr->e);
  memset(arr, 0, sizeof(*arr));
}

}  // extern "C"

extern "C" mf_status mf_flow_return_excess(const mf_csr *g, int32_t *flow) {
  if (!g || (!flow && g->m > 0)) return MF_ERR_INVALID_ARGUMENT;
  mf_status st = mf::validate_csr(g);
  if (st != MF_OK) return st;
  mf::return_excess_to_source(g, flow);
  return MF_OK;
}
