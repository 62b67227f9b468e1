/*
 * oracle.h - CPU oracle for the push-relabel max-flow path.  TEST INFRASTRUCTURE ONLY:
 * nothing in the product (wbpr_amd/) may link or call this code; it is the checker used by
 * tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg.
 *
 *   oracle_serial_maxflow  restatement of the reference CPU solver maxflow-serial
 *                          (maxflow-serial/graph.cpp:316-619), quirks included
 *   oracle_dinic_maxflow   independent exact max flow (Dinic's blocking flows)
 *   oracle_certify         capacity / conservation / value / min-cut certificate of a flow
 *   oracle_rcsr_build, oracle_bcsr_build   the reference's residual-layout builders
 *
 * All arrays are CSR: offsets[n+1], dst[m], cap[m] (int32), 0-based vertex ids.
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Returns the max-flow value as maxflow-serial computes it, or -1 when checkPath() finds no
 * s -> t path (the reference then prints "No path from source to sink" and no value).
 * ops_out (optional) receives the number of push + relabel operations. */
long long oracle_serial_maxflow(int n, int m, const int *offsets, const int *dst, const int *cap,
                                int s, int t, long long *ops_out);

/* Exact maximum flow; flow_out (optional, m entries) receives a maximum flow. */
long long oracle_dinic_maxflow(int n, int m, const int *offsets, const int *dst, const int *cap,
                               int s, int t, int *flow_out);

/* Certificate of a flow with the claimed value.  Returns 0 when every check passes, else:
 *   1 capacity violated (flow < 0 or > cap)
 *   2 conservation violated at some v not in {s, t} (or negative excess if preflow != 0)
 *   3 net inflow at t differs from value
 *   4 t reachable from s in the residual graph, or cut capacity differs from value
 * With preflow != 0 the flow may leave non-negative excess at vertices; the cut is then taken
 * on the sink side (vertices that reach t in the residual graph).
 * cut_out (optional) receives the capacity of the certificate cut. */
int oracle_certify(int n, int m, const int *offsets, const int *dst, const int *cap, int s, int t,
                   const int *flow, long long value, int preflow, long long *cut_out);

/* Reference residual-layout builders (see layouts.c). */
void oracle_rcsr_build(int n, int m, const int *offsets, const int *dst, int *roffsets, int *rdst,
                       int *flow_index);
void oracle_bcsr_build(int n, int m, const int *arc_u, const int *arc_v, const int *arc_cap,
                       int *offsets, int *dst, int *residual, int *backward_idx);

#ifdef __cplusplus
}
#endif
#endif
