// Dumps what the reference's own host code builds from a graph file: the CSRGraph its loaders
// produce and the ResidualGraph its buildFromCSRGraph derives.  TEST INFRASTRUCTURE ONLY.
//
// This driver is compiled by oracle/Makefile together with one reference tree's unmodified
// src/graph.cpp and src/mmio.c (where they lie under the reference checkout), once per layout:
//   REF_RCSR   maxflow-cuda   (graph.cpp:28-67 txt, 151-214 DIMACS, 292-314 binary, 317-402 RCSR)
//   REF_BCSR   maxflow-bcsr   (graph.cpp:151-219 DIMACS with a reverse slot per arc, 322-365 BCSR)
//   REF_BCSR2  maxflow-bcsr2  (graph.cpp:322-365: interleaved {residual, reverse slot} pairs)
// The tests compare these arrays with the library's loaders and layout builders.
//
//   ref_layouts_<layout> <file type 0 txt | 1 binary | 2 DIMACS> <input> <output>
//
// Output: one line per array, "name count v0 v1 ...", in the order written below.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "graph.h"

static void dump(FILE *f, const char *name, const int *a, long long n) {
  fprintf(f, "%s %lld", name, n);
  for (long long i = 0; i < n; ++i) fprintf(f, " %d", a[i]);
  fputc('\n', f);
}

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <0 txt | 1 binary | 2 DIMACS> <input> <output>\n", argv[0]);
    return 2;
  }
  const int type = atoi(argv[1]);
  CSRGraph g;
  g.num_nodes = g.num_edges = g.num_edges_processed = 0;
  g.source_node = g.sink_node = -1;
  if (type == 0) g.buildFromTxtFile(argv[2]);
  else if (type == 1) g.loadFromBinary(argv[2]);
  else if (type == 2) g.buildFromDIMACSFile(argv[2]);
  else return 2;
  FILE *f = fopen(argv[3], "w");
  if (!f) return 1;
  const int head[5] = {g.num_nodes, g.num_edges, g.num_edges_processed, g.source_node,
                       g.sink_node};
  dump(f, "csr_header", head, 5);
  dump(f, "csr_offsets", g.offsets.data(), (long long)g.offsets.size());
  dump(f, "csr_destinations", g.destinations.data(), (long long)g.destinations.size());
  dump(f, "csr_capacities", g.capacities.data(), (long long)g.capacities.size());
  ResidualGraph r;
  r.buildFromCSRGraph(g);
  const long long m = r.num_edges;
  dump(f, "offsets", r.offsets, r.num_nodes + 1);
  dump(f, "destinations", r.destinations, m);
  dump(f, "capacities", r.capacities, m);
#if defined(REF_RCSR)
  dump(f, "forward_flows", r.forward_flows, m);
  dump(f, "backward_flows", r.backward_flows, m);
  dump(f, "roffsets", r.roffsets, r.num_nodes + 1);
  dump(f, "rdestinations", r.rdestinations, m);
  dump(f, "flow_index", r.flow_index, m);
#elif defined(REF_BCSR)
  dump(f, "forward_flows", r.forward_flows, m);
  dump(f, "backward_idx", r.backward_idx, m);
#elif defined(REF_BCSR2)
  dump(f, "forward_flows", r.forward_flows, 2 * m);
#else
#error "define REF_RCSR, REF_BCSR or REF_BCSR2"
#endif
  fclose(f);
  return 0;
}
