/* oracle.h — CPU oracles for the GST DPQ solve. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and only as the checker or the reported CPU baseline; the product
 * (libgst.so and everything above it) never links or calls it.
 *
 * Functions:
 *   orc_prunedpp   restatement of PrunedDP++ (code/PrunedDP++/include/CPUNONHOP.h:499-967)
 *   orc_dprunedpp  restatement of D-PrunedDP++ (code/D-PrunedDP++/include/CPUHOP.h:107-288)
 *   orc_dpbf       independent unpruned DPBF (Ding et al., ICDE 2007), best-first over (v,p)
 *   orc_dpbf_hop   independent layered hop DP: subsets by popcount, heights in order
 *   orc_group_sssp multi-source Dijkstra from every query group (distances only)
 *   orc_*_batch    one query per worker thread (the CPU baseline of the benchmark)
 *
 * Graph: symmetric CSR beg[n+1] (int64), adj/w[m] (int32, adjacency sorted by neighbour).
 * Groups: group g = verts[offs[g] .. offs[g+1]). Query: k group ids (duplicates dropped,
 * first occurrence kept, as the reference's unordered_set does).
 * Status: 0 = solved, 1 = no feasible tree, <0 = bad input (k > 14 etc.).
 */
#ifndef GST_ORACLE_H
#define GST_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n;
  const int64_t *beg;
  const int32_t *adj;
  const int32_t *w;
  int32_t n_groups;
  const int64_t *goffs;
  const int32_t *gverts;
} orc_graph;

typedef struct {
  int32_t status;     /* 0 solved, 1 no solution, <0 error */
  int32_t dp_cost;    /* DP value of the returned root state (100000 if none) */
  int32_t tree_cost;  /* reference-reported cost: sum of the distinct edges of the
                         rebuilt tree, 100000 if the tree is infeasible */
  int32_t n_edges;    /* edges written to edges[] (u,v,w triples), capped by edge_cap */
  int32_t root, root_d;
  int64_t counts;     /* states ever assigned a finite cost */
  int64_t process;    /* states popped */
  double t_search_s;  /* search-loop time (reference timing scope) */
  double t_total_s;   /* whole call */
} orc_result;

#define ORC_INF 100000

int orc_prunedpp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *edges,
                 int32_t edge_cap, orc_result *res);
int orc_dprunedpp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t D,
                  int32_t *edges, int32_t edge_cap, orc_result *res);
int orc_dpbf(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *edges,
             int32_t edge_cap, orc_result *res);
int orc_dpbf_hop(const orc_graph *g, const int32_t *gids, int32_t k, int32_t D,
                 orc_result *res);
/* dist[i*n + v] = distance from group gids[i] to v (ORC_INF if unreachable) */
int orc_group_sssp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *dist);

/* Solve n_queries queries (row-major gids[q*k+j]) with n_threads workers, one query per
 * worker at a time. algo: 0 = PrunedDP++, 1 = D-PrunedDP++ (uses D), 2 = DPBF.
 * Stops starting new queries once time_budget_s has elapsed (<= 0: no budget); returns
 * the number of queries solved (a prefix is not guaranteed: results[q].status = -99
 * marks queries not run). */
int orc_batch(const orc_graph *g, const int32_t *gids, int32_t n_queries, int32_t k,
              int32_t D, int32_t algo, int32_t n_threads, double time_budget_s,
              orc_result *results);

#ifdef __cplusplus
}
#endif
#endif
