// Host-only driver for the loaders, writers, generators, terminal transforms and layout builders
// of libmf, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitized.py (g++, host code only: host_graph.cpp).  argv[1]: a directory of
// input files, well-formed and malformed; every loader reads every file.  argv[2]: a scratch
// directory for the round-trip files.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <dirent.h>
#include "mf.h"

static void layouts(const mf_host_graph *g) {
  mf_csr c{g->n, g->m, g->offsets, g->dst, g->cap, g->s < 0 ? 0 : g->s, g->t < 0 ? (g->n > 1 ? 1 : 0) : g->t};
  for (int l = 0; l < 3; ++l) {
    mf_layout_arrays a;
    memset(&a, 0, sizeof(a));
    mf_status s = mf_build_layout(&c, (mf_layout)l, &a);
    if (s == MF_OK) mf_layout_free(&a);
  }
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const char *dir = argv[1];
  DIR *d = opendir(dir);
  if (!d) return 3;
  struct dirent *e;
  int files = 0, ok = 0;
  while ((e = readdir(d))) {
    if (e->d_name[0] == '.') continue;
    std::string p = std::string(dir) + "/" + e->d_name;
    typedef mf_status (*L)(const char *, mf_host_graph **);
    L fs[] = {mf_load_dimacs, mf_load_snap_txt, mf_load_snap_bin, mf_load_konect_bipartite};
    for (L f : fs) {
      mf_host_graph *g = nullptr;
      ++files;
      if (f(p.c_str(), &g) == MF_OK && g) {
        ++ok;
        if (g->n > 1 && g->n < 2000000) {
          layouts(g);
          mf_csr c{g->n, g->m, g->offsets, g->dst, g->cap, g->s < 0 ? 0 : g->s, g->t < 0 ? 1 : g->t};
          const std::string tb = std::string(argv[2]) + "/rt.bin", td = std::string(argv[2]) + "/rt.max";
          if (mf_save_snap_bin(tb.c_str(), &c) == MF_OK) {
            mf_host_graph *h = nullptr;
            if (mf_load_snap_bin(tb.c_str(), &h) == MF_OK) mf_host_graph_free(h);
            // truncated copies of the binary
            FILE *in = fopen(tb.c_str(), "rb");
            std::vector<char> buf;
            int ch;
            while ((ch = fgetc(in)) != EOF) buf.push_back((char)ch);
            fclose(in);
            for (size_t cut : {(size_t)0, (size_t)3, buf.size() / 2, buf.size() - 1}) {
              if (cut > buf.size()) continue;
              const std::string tc = std::string(argv[2]) + "/cut.bin";
              FILE *o = fopen(tc.c_str(), "wb");
              fwrite(buf.data(), 1, cut, o);
              fclose(o);
              mf_host_graph *h2 = nullptr;
              if (mf_load_snap_bin(tc.c_str(), &h2) == MF_OK) mf_host_graph_free(h2);
            }
          }
          if (mf_save_dimacs(td.c_str(), &c) == MF_OK) {
            mf_host_graph *h = nullptr;
            if (mf_load_dimacs(td.c_str(), &h) == MF_OK) mf_host_graph_free(h);
          }
          int32_t srcs[2] = {0, g->n - 1}, snks[1] = {g->n / 2};
          mf_host_graph *st = nullptr;
          if (mf_super_terminals(&c, srcs, 2, snks, 1, &st) == MF_OK) mf_host_graph_free(st);
          mf_host_graph *bp = nullptr;
          int32_t so[3], si[3];
          if (mf_bfs_terminal_pairs(&c, 3, 7, so, si, &bp) == MF_OK) mf_host_graph_free(bp);
          std::vector<int32_t> flow((size_t)g->m, 0);
          for (int64_t a = 0; a < g->m; ++a) flow[(size_t)a] = g->cap ? g->cap[a] / 2 : 0;
          (void)mf_flow_return_excess(&c, flow.data());
        }
        mf_host_graph_free(g);
      }
    }
  }
  closedir(d);
  mf_host_graph *g = nullptr;
  if (mf_gen_rmat(12, 8, 0.57, 0.19, 0.19, 1000, 3, &g) == MF_OK) { layouts(g); mf_host_graph_free(g); }
  if (mf_gen_genrmf(8, 5, 1, 100, 2, &g) == MF_OK) { layouts(g); mf_host_graph_free(g); }
  if (mf_gen_washington_rlg(16, 16, 100, 2, &g) == MF_OK) { layouts(g); mf_host_graph_free(g); }
  if (mf_gen_bipartite(300, 200, 4.0, 2, &g) == MF_OK) { layouts(g); mf_host_graph_free(g); }
  printf("loader calls %d, parsed %d\n", files, ok);
  return 0;
}
