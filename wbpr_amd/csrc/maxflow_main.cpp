// maxflow: command-line driver with the reference's interface (maxflow-cuda/src/main.cu:7-240).
//
//   maxflow [-h] [-v 0|1|2] [-f file] [-s source] [-t sink] [-a 0|1] [-l 0|1|2] [-k cycles] [-c]
//
// As in the reference, -f loads the file at once with the file type given so far, so -v must
// come before -f, and -s/-t are ignored for DIMACS input (source and sink come from its 'n'
// lines).  Output keeps the lines the reference's scripts parse (bash_run.py:18-22):
//   "Total kernel time: %.6f ms"
//   "The maximum flow value of this flow network as calculated by the parallel implementation is %d, %d"
// Extra options of this build: -l residual layout (0 RCSR, 1 BCSR, 2 BCSR2; default 1),
// -k push-relabel cycles per global relabel, -c certificate check, -o DIMACS flow output file
// (the s/f lines of OutputFlow, network/DIMACS/washington/washington.c:1333-1399), -x minimum-cut
// output file.
#include <getopt.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mf.h"

static void usage(const char *prog) {
  printf("Usage: %s [-h] [-v file type] [-f filename] [-s source] [-t sink]\n", prog);
  printf("Options:\n");
  printf("\t-h\t\tShow this help message and exit\n");
  printf("\t-v\t\tSpecify which kind of file to load\n");
  printf("\t\t\t\t0: SNAP txt file (default)\n");
  printf("\t\t\t\t1: SNAP binary file\n");
  printf("\t\t\t\t2: DIMACS file\n");
  printf("\t\t\t\t3: KONECT bipartite edge list (matching network)\n");
  printf("\t-f filename\tSpecify the file path (binary or txt)\n");
  printf("\t-s source\tSource node\n");
  printf("\t-t sink\t\tSink node\n");
  printf("\t-a algorithm\tSpecify the algorithm to use\n");
  printf("\t\t\t\t0: Thread-centric push-relabel (default)\n");
  printf("\t\t\t\t1: Vertex-centric push-relabel\n");
  printf("\t-l layout\tResidual layout: 0 RCSR, 1 BCSR (default), 2 BCSR2\n");
  printf("\t-k cycles\tPush-relabel cycles between global relabels (default automatic)\n");
  printf("\t-c\t\tCheck the result: capacity, conservation and min-cut certificate\n");
  printf("\t-o file\t\tWrite the flow in DIMACS solution format (s/f lines)\n");
  printf("\t-x file\t\tWrite a minimum cut: 'c cut <capacity>', then 'n <vertex> s|t' per vertex\n");
  printf("\t-w\t\tPrint the busy cycles of every wave of the push-relabel launches\n");
  printf("\t-m pairs\tSNAP input without -s/-t: super-source and super-sink over BFS-chosen pairs\n");
}

// Graph statistics in the format of CSRGraph::printGraphStatus (maxflow-cuda/src/graph.cpp:668-683).
static void print_graph_status(const mf_host_graph *g) {
  double sum = 0, sq = 0;
  int maxd = 0, maxv = 0, mind = g->n, minv = 0;
  for (int v = 0; v < g->n; ++v) {
    int d = g->offsets[v + 1] - g->offsets[v];
    sum += d;
    if (d > maxd) {
      maxd = d;
      maxv = v;
    }
    if (d < mind) {
      mind = d;
      minv = v;
    }
  }
  double avg = sum / g->n;
  for (int v = 0; v < g->n; ++v) {
    double d = g->offsets[v + 1] - g->offsets[v] - avg;
    sq += d * d;
  }
  printf("---------<< Graph Status >>---------\n");
  printf("Number of nodes: %d\n", g->n);
  printf("Number of edges: %lld\n", (long long)g->m);
  printf("Average degree: %.2f\n", avg);
  printf("Standard deviation of degree: %.2f\n", std::sqrt(sq / g->n));
  printf("Maximum degree: %d (Node %d)\n", maxd, maxv);
  printf("Minimum degree: %d (Node %d)\n", mind, minv);
  printf("Source node: %d, Sink node: %d\n", g->s, g->t);
}

// Capacity, conservation and cut checks of a flow (returns 0 when all hold).
static int check_certificate(const mf_csr *c, const int32_t *flow, long long value,
                             const uint8_t *sink_side) {
  std::vector<long long> net((size_t)c->n, 0);
  for (int u = 0; u < c->n; ++u)
    for (int i = c->offsets[u]; i < c->offsets[u + 1]; ++i) {
      int cap = c->cap ? c->cap[i] : 1;
      if (flow[i] < 0 || flow[i] > cap) {
        printf("Certificate: capacity violated on arc %d\n", i);
        return 1;
      }
      net[(size_t)u] -= flow[i];
      net[(size_t)c->dst[i]] += flow[i];
    }
  for (int v = 0; v < c->n; ++v)
    if (v != c->s && v != c->t && net[(size_t)v] != 0) {
      printf("Certificate: conservation violated at vertex %d (%lld)\n", v, net[(size_t)v]);
      return 1;
    }
  if (net[(size_t)c->t] != value) {
    printf("Certificate: net inflow at sink %lld != value %lld\n", net[(size_t)c->t], value);
    return 1;
  }
  long long cut = 0;
  for (int u = 0; u < c->n; ++u)
    for (int i = c->offsets[u]; i < c->offsets[u + 1]; ++i)
      if (!sink_side[u] && sink_side[c->dst[i]]) cut += c->cap ? c->cap[i] : 1;
  if (sink_side[c->s] || !sink_side[c->t] || cut != value) {
    printf("Certificate: cut capacity %lld != value %lld\n", cut, value);
    return 1;
  }
  printf("Certificate: ok (cut capacity %lld)\n", cut);
  return 0;
}

int main(int argc, char **argv) {
  int ch, file_type = 0, algo_type = 0, layout = 1, cycles = 0, check = 0, workload = 0, pairs = 0;
  int source = -1, sink = -1;
  const char *filename = nullptr, *outfile = nullptr, *cutfile = nullptr;
  mf_host_graph *g = nullptr;
  while ((ch = getopt(argc, argv, "hv:f:s:t:a:l:k:co:x:wm:")) != -1) {
    switch (ch) {
      case 'h':
        usage(argv[0]);
        return 0;
      case 'v':
        file_type = atoi(optarg);
        break;
      case 'f': {
        filename = optarg;
        mf_status st;
        if (g) mf_host_graph_free(g);
        g = nullptr;
        switch (file_type) {
          case 0:
            printf("Loading from txt file: %s\n", filename);
            st = mf_load_snap_txt(filename, &g);
            break;
          case 1:
            printf("Loading from binary file: %s\n", filename);
            st = mf_load_snap_bin(filename, &g);
            break;
          case 2:
            printf("Loading from DIMACS file: %s\n", filename);
            st = mf_load_dimacs(filename, &g);
            break;
          case 3:
            printf("Loading from KONECT bipartite file: %s\n", filename);
            st = mf_load_konect_bipartite(filename, &g);
            break;
          default:
            printf("Invalid file type\n");
            return 1;
        }
        if (st != MF_OK) {
          fprintf(stderr, "Error loading %s: %s\n", filename, mf_status_string(st));
          return 1;
        }
        break;
      }
      case 's':
        if (file_type != 2) source = atoi(optarg);
        break;
      case 't':
        if (file_type != 2) sink = atoi(optarg);
        break;
      case 'a':
        algo_type = atoi(optarg);
        break;
      case 'l':
        layout = atoi(optarg);
        break;
      case 'k':
        cycles = atoi(optarg);
        break;
      case 'c':
        check = 1;
        break;
      case 'o':
        outfile = optarg;
        break;
      case 'x':
        cutfile = optarg;
        break;
      case 'w':
        workload = 1;
        break;
      case 'm':
        pairs = atoi(optarg);
        break;
      default:
        printf("Invalid option\n");
        return 1;
    }
  }
  if (!g) {
    printf("No input file (-f)\n");
    return 1;
  }
  if (file_type < 2) {
    g->s = source;
    g->t = sink;
  }
  if (pairs > 0 && file_type < 2) {
    mf_csr v = mf_host_graph_view(g);
    mf_host_graph *h = nullptr;
    const mf_status st = mf_bfs_terminal_pairs(&v, pairs, 1, nullptr, nullptr, &h);
    if (st != MF_OK) {
      printf("Cannot choose %d source/sink pairs: %s\n", pairs, mf_status_string(st));
      return 1;
    }
    mf_host_graph_free(g);
    g = h;
    printf("Super-source %d and super-sink %d over %d BFS-chosen pairs\n", g->s, g->t, pairs);
  }
  print_graph_status(g);
  if (algo_type == 0) {
    printf("Using thread-centric push-relabel algorithm\n");
  } else if (algo_type == 1) {
    printf("Using vertex-centric push-relabel algorithm\n");
  } else {
    printf("Invalid algorithm type\n");
    return 1;
  }
  if (layout < 0 || layout > 2) {
    printf("Invalid layout\n");
    return 1;
  }
  mf_csr c = mf_host_graph_view(g);
  if (c.s < 0 || c.s >= c.n || c.t < 0 || c.t >= c.n || c.s == c.t) {
    printf("Invalid source/sink: %d %d\n", c.s, c.t);
    return 1;
  }
  mf_context *ctx = nullptr;
  mf_status st = mf_context_create(0, &ctx);
  if (st != MF_OK) {
    fprintf(stderr, "GPU setup failed: %s\n", mf_status_string(st));
    return 1;
  }
  char name[300];
  int32_t ncu = 0;
  mf_device_info(ctx, &ncu, name, sizeof(name));
  printf("GPU Device: %s, %d CUs\n", name, ncu);
  mf_options o;
  mf_options_default(&o);
  o.cycles_per_relabel = cycles;
  o.record_workload = workload;
  mf_graph *dg = nullptr;
  st = mf_graph_upload(ctx, &c, (mf_layout)layout, &dg);
  mf_result r;
  memset(&r, 0, sizeof(r));
  if (st == MF_OK) st = mf_graph_solve(ctx, dg, (mf_algo)algo_type, &o, &r);
  if (st != MF_OK) {
    fprintf(stderr, "Solve failed: %s\n", mf_status_string(st));
    if (dg) mf_graph_destroy(ctx, dg);
    mf_context_destroy(ctx);
    return 1;
  }
  printf("Global relabels: %lld, push-relabel cycles: %lld, pushes: %lld, relabels: %lld\n",
         (long long)r.rounds, (long long)r.cycles, (long long)r.pushes, (long long)r.relabels);
  printf("Global relabel time: %.6f ms\n", r.relabel_ms);
  printf("Solve time: %.6f ms\n", r.solve_ms);
  printf("Total kernel time: %.6f ms\n", r.kernel_ms);
  printf("The maximum flow value of this flow network as calculated by the parallel "
         "implementation is %lld, %lld\n",
         (long long)r.value, (long long)r.value);
  int rc = 0;
  if (workload) {  // the WORKLOAD report of maxflow-cuda/src/push_relabel.cu:202-210
    int64_t nw = 0;
    std::vector<uint64_t> cyc;
    if (mf_graph_read_workload(ctx, dg, nullptr, nullptr, 0, &nw) == MF_OK && nw > 0) {
      cyc.resize((size_t)nw);
      if (mf_graph_read_workload(ctx, dg, cyc.data(), nullptr, nw, &nw) != MF_OK) nw = 0;
    }
    printf("------------<< Workload Information >>------------\n");
    printf("#warps: %lld\n", (long long)nw);
    printf("Warp execution time:\n");
    for (int64_t i = 0; i < nw; ++i) printf("%llu ", (unsigned long long)cyc[(size_t)i]);
    printf("\n");
  }
  if (check || outfile || cutfile) {
    std::vector<int32_t> flow((size_t)c.m);
    std::vector<uint8_t> side((size_t)c.n);
    if (mf_graph_read_flow(ctx, dg, flow.data(), 1) != MF_OK ||
        mf_graph_read_cut(ctx, dg, side.data()) != MF_OK) {
      fprintf(stderr, "Flow read-back failed\n");
      rc = 1;
    } else {
      if (check) rc = check_certificate(&c, flow.data(), r.value, side.data());
      if (outfile) {
        FILE *f = fopen(outfile, "w");
        if (!f) {
          fprintf(stderr, "Cannot write %s\n", outfile);
          rc = 1;
        } else {
          fprintf(f, "s %lld\n", (long long)r.value);
          for (int u = 0; u < c.n; ++u)
            for (int i = c.offsets[u]; i < c.offsets[u + 1]; ++i)
              if (flow[(size_t)i] > 0) fprintf(f, "f %d %d %d\n", u + 1, c.dst[i] + 1, flow[(size_t)i]);
          fclose(f);
        }
      }
      if (cutfile) {  // source side: the vertices that cannot reach t in the residual graph
        FILE *f = fopen(cutfile, "w");
        if (!f) {
          fprintf(stderr, "Cannot write %s\n", cutfile);
          rc = 1;
        } else {
          long long cut = 0;
          for (int u = 0; u < c.n; ++u)
            for (int i = c.offsets[u]; i < c.offsets[u + 1]; ++i)
              if (!side[(size_t)u] && side[(size_t)c.dst[i]]) cut += c.cap ? c.cap[i] : 1;
          fprintf(f, "c cut %lld\n", cut);
          for (int u = 0; u < c.n; ++u) fprintf(f, "n %d %c\n", u + 1, side[(size_t)u] ? 't' : 's');
          fclose(f);
        }
      }
    }
  }
  mf_graph_destroy(ctx, dg);
  mf_context_destroy(ctx);
  mf_host_graph_free(g);
  return rc;
}
