/*
 * serial.c - restatement of the reference CPU solver maxflow-serial, for use as the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows maxflow-serial/graph.cpp function by function, with the reference's behaviour kept,
 * including its quirks (so values match the reference binary on the same input):
 *   - residual graph (graph.cpp:316-401): forward residual = cap, backward residual = 0,
 *     reverse CSR in (tail, arc) order, flow_index[k] = FIRST arc of out(v) with dst == u
 *     (parallel arcs alias to one arc);
 *   - preflow (graph.cpp:463-479): h(s) = n; for every arc out of s the head's excess is
 *     OVERWRITTEN with the capacity (parallel s -> v arcs undercount);
 *   - findActiveNode (graph.cpp:527-540): minimum height among v != s,t with e > 0 and h < n,
 *     lowest id on ties;
 *   - push (graph.cpp:482-518): first forward arc with h(v) == h(w)+1 and non-zero push, else the
 *     first such reverse arc; relabel (graph.cpp:521-524): h += 1;
 *   - maxflow (graph.cpp:555-589): checkPath() first (DFS over forward arcs, graph.cpp:593-619);
 *     value = e(t).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int check_path(int n, const int *off, const int *dst, int s, int t) {
  /* graph.cpp:593-619 (stack-based search over forward arcs) */
  char *visited = (char *)calloc((size_t)n, 1);
  int *stack = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  int top = 0, found = 0;
  stack[top++] = s;
  visited[s] = 1;
  while (top > 0 && !found) {
    int u = stack[--top];
    for (int i = off[u]; i < off[u + 1]; ++i) {
      int v = dst[i];
      if (!visited[v]) {
        if (v == t) {
          found = 1;
          break;
        }
        stack[top++] = v;
        visited[v] = 1;
      }
    }
  }
  if (!found) found = visited[t];
  free(visited);
  free(stack);
  return found;
}

long long oracle_serial_maxflow(int n, int m, const int *off, const int *dst, const int *cap, int s,
                                int t, long long *ops_out) {
  long long ops = 0;
  if (!check_path(n, off, dst, s, t)) {
    if (ops_out) *ops_out = 0;
    return -1;
  }
  int *ff = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int *bf = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int *roff = (int *)calloc((size_t)n + 1, sizeof(int));
  int *rdst = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int *fidx = (int *)malloc(sizeof(int) * (size_t)(m + 1));
  int *h = (int *)calloc((size_t)n, sizeof(int));
  int *e = (int *)calloc((size_t)n, sizeof(int));
  int *cnt = (int *)calloc((size_t)n, sizeof(int));
  /* buildFromCSRGraph, graph.cpp:343-398 */
  for (int i = 0; i < m; ++i) {
    ff[i] = cap[i];
    bf[i] = 0;
  }
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
  /* preflow, graph.cpp:463-479 */
  h[s] = n;
  for (int i = off[s]; i < off[s + 1]; ++i) {
    e[dst[i]] = cap[i];
    ff[i] = 0;
    bf[i] = cap[i];
  }
  /* main loop, graph.cpp:572-582 */
  for (;;) {
    /* findActiveNode, graph.cpp:527-540 */
    int best = n, v = -1;
    for (int i = 0; i < n; ++i)
      if (e[i] > 0 && i != s && i != t && h[i] < best) {
        best = h[i];
        v = i;
      }
    if (v < 0) break;
    ++ops;
    /* push, graph.cpp:482-518 */
    int pushed = 0;
    for (int i = off[v]; i < off[v + 1] && !pushed; ++i) {
      int w = dst[i];
      if (h[v] == h[w] + 1) {
        int f = e[v] < ff[i] ? e[v] : ff[i];
        if (f == 0) continue;
        ff[i] -= f;
        bf[i] += f;
        e[v] -= f;
        e[w] += f;
        pushed = 1;
      }
    }
    for (int k = roff[v]; k < roff[v + 1] && !pushed; ++k) {
      int w = rdst[k];
      if (h[v] == h[w] + 1) {
        int pi = fidx[k];
        int f = e[v] < bf[pi] ? e[v] : bf[pi];
        if (f == 0) continue;
        bf[pi] -= f;
        ff[pi] += f;
        e[v] -= f;
        e[w] += f;
        pushed = 1;
      }
    }
    /* relabel, graph.cpp:521-524 */
    if (!pushed) h[v] += 1;
  }
  long long value = e[t];
  free(ff);
  free(bf);
  free(roff);
  free(rdst);
  free(fidx);
  free(h);
  free(e);
  free(cnt);
  if (ops_out) *ops_out = ops;
  return value;
}
