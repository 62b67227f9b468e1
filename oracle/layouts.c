/*
 * layouts.c - restatements of the reference residual-layout builders, for layout tests.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   oracle_rcsr_build: ResidualGraph::buildFromCSRGraph of maxflow-cuda/src/graph.cpp:352-398
 *     (reverse CSR by in-degree counting, reverse entries in (tail, arc) order, flow_index[k] =
 *     first arc of out(v) with destination u, found by linear search).
 *   oracle_bcsr_build: the BCSR DIMACS loader (maxflow-bcsr/src/graph.cpp:187-212: each arc
 *     u -> v appends (v, cap) to u's list and (u, 0) to v's list, in file order) followed by
 *     ResidualGraph::buildFromCSRGraph (maxflow-bcsr/src/graph.cpp:350-362: backward index of
 *     slot i = first slot of dst(i) whose destination is u, by linear search).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void oracle_rcsr_build(int n, int m, const int *off, const int *dst, int *roff, int *rdst,
                       int *fidx) {
  int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
  memset(roff, 0, sizeof(int) * ((size_t)n + 1));
  for (int u = 0; u < n; ++u)
    for (int j = off[u]; j < off[u + 1]; ++j) cnt[dst[j]]++;
  for (int v = 1; v <= n; ++v) roff[v] = roff[v - 1] + cnt[v - 1];
  memset(cnt, 0, sizeof(int) * (size_t)n);
  for (int u = 0; u < n; ++u)
    for (int j = off[u]; j < off[u + 1]; ++j) {
      int d = dst[j];
      rdst[roff[d] + cnt[d]++] = u;
    }
  for (int u = 0; u < n; ++u)
    for (int k = roff[u]; k < roff[u + 1]; ++k) {
      int v = rdst[k];
      fidx[k] = -1;
      for (int j = off[v]; j < off[v + 1]; ++j)
        if (dst[j] == u) {
          fidx[k] = j;
          break;
        }
    }
  (void)m;
  free(cnt);
}

void oracle_bcsr_build(int n, int m, const int *au, const int *av, const int *acap, int *off,
                       int *dst, int *res, int *bidx) {
  /* adjacency lists in file order */
  int *deg = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int i = 0; i < m; ++i) {
    deg[au[i]]++;
    deg[av[i]]++;
  }
  off[0] = 0;
  for (int v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
  int *pos = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  memcpy(pos, off, sizeof(int) * (size_t)n);
  for (int i = 0; i < m; ++i) {
    int a = pos[au[i]]++;
    dst[a] = av[i];
    res[a] = acap[i];
    int b = pos[av[i]]++;
    dst[b] = au[i];
    res[b] = 0;
  }
  for (int u = 0; u < n; ++u)
    for (int i = off[u]; i < off[u + 1]; ++i) {
      int v = dst[i];
      bidx[i] = -1;
      for (int j = off[v]; j < off[v + 1]; ++j)
        if (dst[j] == u) {
          bidx[i] = j;
          break;
        }
    }
  free(deg);
  free(pos);
}
