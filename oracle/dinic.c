/*
 * dinic.c - exact maximum flow by Dinic's algorithm (BFS level graph + blocking flow with
 * current-arc pointers), independent of push-relabel.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Residual arcs are stored in pairs (2i forward, 2i+1 backward) so the reverse of arc a is a^1.
 * Capacities are int32, sums int64.  Parallel and antiparallel arcs are handled exactly.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

long long oracle_dinic_maxflow(int n, int m, const int *off, const int *dst, const int *cap, int s,
                               int t, int *flow_out) {
  if (s == t) return 0;
  const long long M = 2LL * m;
  int *head = (int *)malloc(sizeof(int) * (size_t)(M + 1));
  int *res = (int *)malloc(sizeof(int) * (size_t)(M + 1));
  int *first = (int *)calloc((size_t)n + 1, sizeof(int));
  int *adj = (int *)malloc(sizeof(int) * (size_t)(M + 1)); /* arcs grouped by tail */
  int *level = (int *)malloc(sizeof(int) * (size_t)n);
  int *it = (int *)malloc(sizeof(int) * (size_t)n);
  int *queue = (int *)malloc(sizeof(int) * (size_t)n);
  int *path = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  for (int u = 0; u < n; ++u)
    for (int i = off[u]; i < off[u + 1]; ++i) {
      head[2 * (long long)i] = dst[i];
      head[2 * (long long)i + 1] = u;
      res[2 * (long long)i] = cap[i];
      res[2 * (long long)i + 1] = 0;
      first[u + 1]++;
      first[dst[i] + 1]++;
    }
  for (int v = 0; v < n; ++v) first[v + 1] += first[v];
  {
    int *pos = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(pos, first, sizeof(int) * (size_t)n);
    for (int u = 0; u < n; ++u)
      for (int i = off[u]; i < off[u + 1]; ++i) {
        adj[pos[u]++] = 2 * i;
        adj[pos[dst[i]]++] = 2 * i + 1;
      }
    free(pos);
  }
  long long total = 0;
  for (;;) {
    /* BFS levels from s */
    for (int v = 0; v < n; ++v) level[v] = -1;
    int qh = 0, qt = 0;
    level[s] = 0;
    queue[qt++] = s;
    while (qh < qt) {
      int u = queue[qh++];
      for (int k = first[u]; k < first[u + 1]; ++k) {
        int a = adj[k];
        int w = head[a];
        if (res[a] > 0 && level[w] < 0) {
          level[w] = level[u] + 1;
          queue[qt++] = w;
        }
      }
    }
    if (level[t] < 0) break;
    for (int v = 0; v < n; ++v) it[v] = first[v];
    /* blocking flow: advance / retreat / augment along the level graph */
    int len = 0, v = s;
    for (;;) {
      if (v == t) {
        int b = res[path[0]];
        for (int q = 1; q < len; ++q)
          if (res[path[q]] < b) b = res[path[q]];
        int cut = -1;
        for (int q = 0; q < len; ++q) {
          res[path[q]] -= b;
          res[path[q] ^ 1] += b;
          if (cut < 0 && res[path[q]] == 0) cut = q;
        }
        total += b;
        len = cut;                             /* retreat to the tail of the first saturated arc */
        v = (len == 0) ? s : head[path[len - 1]];
        continue;
      }
      int advanced = 0;
      while (it[v] < first[v + 1]) {
        int a = adj[it[v]];
        int w = head[a];
        if (res[a] > 0 && level[w] == level[v] + 1) {
          path[len++] = a;
          v = w;
          advanced = 1;
          break;
        }
        it[v]++;
      }
      if (advanced) continue;
      /* dead end: remove v from the level graph and retreat */
      level[v] = -1;
      if (len == 0) break;
      len--;
      v = head[path[len] ^ 1];
      it[v]++;
    }
  }
  if (flow_out)
    for (int i = 0; i < m; ++i) flow_out[i] = cap[i] - res[2 * (long long)i];
  free(head);
  free(res);
  free(first);
  free(adj);
  free(level);
  free(it);
  free(queue);
  free(path);
  return total;
}
