/*
 * certify.c - optimality certificate of a flow: capacity, conservation, value and a cut whose
 * capacity equals the value (max-flow min-cut theorem).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int oracle_certify(int n, int m, const int *off, const int *dst, const int *cap, int s, int t,
                   const int *flow, long long value, int preflow, long long *cut_out) {
  long long *net = (long long *)calloc((size_t)n, sizeof(long long));
  int *tail = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int rc = 0;
  for (int u = 0; u < n; ++u)
    for (int i = off[u]; i < off[u + 1]; ++i) {
      tail[i] = u;
      if (flow[i] < 0 || flow[i] > cap[i]) rc = rc ? rc : 1;
      net[u] -= flow[i];
      net[dst[i]] += flow[i];
    }
  for (int v = 0; v < n && !rc; ++v) {
    if (v == s || v == t) continue;
    if (preflow ? net[v] < 0 : net[v] != 0) rc = 2;
  }
  if (!rc && net[t] != value) rc = 3;
  if (!rc) {
    /* residual reachability: from s forwards (flow) or into t backwards (preflow) */
    int *roff = (int *)calloc((size_t)n + 1, sizeof(int));
    int *rarc = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    for (int i = 0; i < m; ++i) roff[dst[i] + 1]++;
    for (int v = 0; v < n; ++v) roff[v + 1] += roff[v];
    int *pos = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(pos, roff, sizeof(int) * (size_t)n);
    for (int i = 0; i < m; ++i) rarc[pos[dst[i]]++] = i;
    char *mark = (char *)calloc((size_t)n, 1);
    int *queue = (int *)malloc(sizeof(int) * (size_t)n);
    int qh = 0, qt = 0;
    const int root = preflow ? t : s;
    mark[root] = 1;
    queue[qt++] = root;
    while (qh < qt) {
      int u = queue[qh++];
      if (!preflow) {
        /* u -> w residual if flow < cap on u->w, or flow > 0 on w->u */
        for (int i = off[u]; i < off[u + 1]; ++i)
          if (flow[i] < cap[i] && !mark[dst[i]]) mark[dst[i]] = 1, queue[qt++] = dst[i];
        for (int k = roff[u]; k < roff[u + 1]; ++k) {
          int i = rarc[k];
          if (flow[i] > 0 && !mark[tail[i]]) mark[tail[i]] = 1, queue[qt++] = tail[i];
        }
      } else {
        /* w reaches u if w -> u is residual: flow < cap on w->u, or flow > 0 on u->w */
        for (int k = roff[u]; k < roff[u + 1]; ++k) {
          int i = rarc[k];
          if (flow[i] < cap[i] && !mark[tail[i]]) mark[tail[i]] = 1, queue[qt++] = tail[i];
        }
        for (int i = off[u]; i < off[u + 1]; ++i)
          if (flow[i] > 0 && !mark[dst[i]]) mark[dst[i]] = 1, queue[qt++] = dst[i];
      }
    }
    long long cut = 0;
    int bad = preflow ? mark[s] : mark[t];
    for (int i = 0; i < m; ++i) {
      int in_src = preflow ? !mark[tail[i]] : mark[tail[i]];
      int in_dst = preflow ? !mark[dst[i]] : mark[dst[i]];
      if (in_src && !in_dst) cut += cap[i];
    }
    if (preflow) {
      /* a maximum preflow leaves no excess on the sink side */
      for (int v = 0; v < n; ++v)
        if (mark[v] && v != t && v != s && net[v] > 0) bad = 1;
    }
    if (bad || cut != value) rc = 4;
    if (cut_out) *cut_out = cut;
    free(roff);
    free(rarc);
    free(pos);
    free(mark);
    free(queue);
  }
  free(net);
  free(tail);
  return rc;
}
