/* gst.h — C ABI of the MI355X group Steiner tree engine (libgst.so, HIP for gfx950).
 *
 * Drop-in for the DPQ solve of the reference's three GPU programs. The reference has no
 * FFI: its solve is a C++ call per query with a CSR graph, a group table and k group ids
 * (plus D for the hop variant). Each entry point below names the reference interface it
 * replaces:
 *
 *   gst_open    replaces the per-run set-up of the drivers: read_input_graph/read_Group +
 *               toCSR + cudaMallocManaged of the CSR
 *               (code/TrimCDP/include/exp_GPU_nonHop.h:40-94,
 *                code/TrimCDP/include/graph_v_of_v_idealID/csr_graph.hpp:39-128,
 *                code/D-TrimCDP-WB/src/GPUHop.cu:96-151). The graph and the group table
 *               stay resident on the device until gst_close.
 *   gst_solve   replaces DP_GPU (code/TrimCDP/include/DPQ.cuh:63, called at
 *               exp_GPU_nonHop.h:122) when D < 0, and DP_gpu
 *               (code/D-TrimCDP/include/DPQ.cuh:63, called at exp_GPU_Hop.h:103) and the
 *               per-query block of D-TrimCDP-WB (code/D-TrimCDP-WB/src/GPUHop.cu:172-265)
 *               when D >= 0. One query, host-driven rounds (one host sync per round, as the
 *               reference). Returns the optimal weight, the tree and the counters the
 *               reference writes to its CSV (time, states, processed states).
 *   gst_solve_batch  the batched persistent form (no reference counterpart: the
 *               reference solves queries one at a time, exp_GPU_nonHop.h:95-136): many
 *               queries of mixed k and D share the device, device-side phase control,
 *               the round loop captured in a HIP graph.
 *   gst_submit / gst_poll / gst_wait  serving: queries are admitted as batch slots and
 *               DP memory (V*2^k*(D+1)*4 B of arena per query, plus its masks and
 *               worklists: gst_job_bytes) free up; results are identical to gst_solve's.
 *               gst_poll on a query in flight reports its upper and lower bounds
 *               (progressive reporting, code/PrunedDP++/include/CPUNONHOP.h:684-726).
 *
 * Conventions (reference): group ids are 0-based indices into the group table (the .g
 * line "g<n>" is group n-1, read_for_GSTP.h:131-133); duplicate ids in a query are
 * dropped (the reference keeps them in an unordered_set). Weights are integers in
 * [1, 2^20]. Status codes replace exit() and the reference's cost sentinels (TrimCDP
 * returns 0 and D-TrimCDP 100000 when there is no solution, DPQ.cu:495-500,571-574 and
 * code/D-TrimCDP/src/DPQ.cu:229-242); drivers apply those sentinels themselves.
 * Thread safety: one gst_solve/gst_solve_batch at a time per context; gst_submit and
 * gst_poll may be called from any thread.
 */
#ifndef GST_H
#define GST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  GST_OK = 0,
  GST_NO_SOLUTION = 1,  /* some query group cannot be connected to the others */
  GST_PENDING = 2,      /* gst_poll: not finished yet (the result holds the progress so far) */
  GST_TREE_OMITTED = 3, /* cost is the optimum; the tree exceeds GST_MAX_TREE_EDGES and is
                           not returned */
  GST_BAD_INPUT = -1,   /* malformed graph/groups/query, k > GST_MAX_K, D > GST_MAX_D */
  GST_OVER_BUDGET = -2, /* the query's DP memory does not fit the memory budget */
  GST_DEVICE_ERROR = -3,
  GST_INTERNAL = -4,    /* reconstruction failed (never expected) */
  GST_NO_DEVICE = -5,
  GST_COST_RANGE = -6   /* no tree below GST_COST_CEILING was found and some cost reached
                           the saturation range: the optimum, if any, is not representable */
};

#define GST_MAX_K 7
#define GST_MAX_D 15
/* DP costs are int32 and saturate at 2^30 - 1; every optimum below GST_COST_CEILING is
 * exact. A query whose costs reach 2^29 and that finds no tree returns GST_COST_RANGE
 * rather than GST_NO_SOLUTION. */
#define GST_COST_CEILING 0x3fffffff
/* Trees of more edges than this are not returned (status GST_TREE_OMITTED). */
#define GST_MAX_TREE_EDGES (1 << 20)
/* Arc offsets are 32-bit on the device, with room for one chunk past the end. */
#define GST_MAX_ARCS ((int64_t)2147483648 - 33554432)

/* flags */
#define GST_F_STATS 1    /* count finite states after the solve (extra arena scan) */
#define GST_F_NO_TREE 2  /* skip tree reconstruction */
#define GST_F_NO_PRUNE 4 /* disable upper/lower-bound pruning (exact either way) */
/* gst_opts.flags only: keep the caller's vertex numbering on the device. By default
 * gst_open renumbers vertices by descending degree (hubs first, so the rows that most
 * arcs point at share a few cache lines); results are always in the caller's ids. */
#define GST_O_KEEP_ORDER 0x100
/* gst_opts.flags only: build no hub distance index at gst_open. By default gst_open
 * computes the distances from the 64 highest-degree vertices to every group, and every
 * non-hop query starts from the upper bound of the best of these 64 stars (the weight of
 * k shortest paths from one vertex). The index holds 64 x n_groups int32 on the host. */
#define GST_O_NO_HUB_INDEX 0x200

typedef struct {
  int32_t n;              /* vertices */
  int64_t m;              /* directed arcs = 2 x undirected edges; at most GST_MAX_ARCS */
  const int64_t *beg_pos; /* [n+1] */
  const int32_t *adj;     /* [m], each adjacency list ascending */
  const int32_t *w;       /* [m], symmetric, in [1, 2^20] */
} gst_csr;

typedef struct {
  int32_t n_groups;
  const int64_t *offs;  /* [n_groups+1] */
  const int32_t *verts; /* members of group g: verts[offs[g] .. offs[g+1]) */
} gst_groups;

typedef struct {
  int32_t device;          /* HIP device ordinal */
  int32_t degree_split;    /* vertices with degree >= this are grown by a wavefront per
                              edge chunk, the rest by a thread per vertex (0 = 32) */
  int32_t chunk_edges;     /* edges per wavefront work item (0 = 1024) */
  int32_t max_batch;       /* job slots of the batched/served forms (0 = 64) */
  int64_t dp_budget_bytes; /* DP memory budget of the batched/served forms
                              (0 = 75% of free device memory at gst_open) */
  int32_t flags;           /* default flags for every solve */
  int32_t graph_rounds;    /* rounds captured per HIP graph launch (0 = 16) */
  int32_t near_far_delta;  /* near-far step of the non-hop grow frontier: states costlier
                              than the threshold wait until the near frontier drains
                              (<= 0 = off: plain label-correcting rounds) */
  int32_t cohort_max;      /* batched/served forms: queued non-hop queries of one k are
                              grown together through level 1, up to this many at a time
                              (0 = up to 64, as many as fit; 1 = each query on its own) */
  int32_t cohort_min;      /* fewest queued queries such a cohort is formed with (0 = 8;
                              smaller cohorts hold up the other jobs for less gain) */
  int32_t cohort_levels;   /* levels a cohort grows together before its members go on as
                              separate jobs (0 = 1; capped at k-1) */
} gst_opts;

typedef struct {
  const int32_t *gids; /* k group ids */
  int32_t k;
  int32_t D;           /* diameter bound, < 0 = none */
  int32_t flags;
} gst_query;

typedef struct {
  int32_t status;         /* GST_OK / GST_NO_SOLUTION / error */
  int32_t cost;           /* optimal weight (-1 when status != GST_OK) */
  int32_t tree_cost;      /* sum of the distinct edge weights of edges[] */
  int32_t n_edges;
  int32_t *edges;         /* n_edges (u, v, w) triples with u < v; owned by the result */
  int32_t root, root_height;
  int32_t k, D;           /* k after de-duplication */
  int32_t rounds;         /* device rounds executed */
  double t_total_s;       /* submission (or call) to result on the host */
  double t_dp_s;          /* DP loop only: the reference's GPU timing scope (DPQ.cu:511-555) */
  double t_queue_s;       /* served form: time spent waiting for admission */
  double t_grow_s;        /* device time of the grow kernels (HIP events, plain path) */
  double t_merge_s;       /* device time of the merge kernels (HIP events, plain path) */
  int64_t states_finite;  /* GST_F_STATS: states with a finite cost (the CSV "counts") */
  int64_t states_processed; /* frontier states evaluated (the CSV "process_num"); level 1
                               of a cohort member counts each state it reached once */
  int64_t grow_vertices, grow_edges, grow_relax, grow_pushes; /* grow-kernel work counters */
  int64_t arena_bytes;    /* DP arena size V*2^k*(D+1)*4 */
  int64_t level_edges[8]; /* grow work per DP level (subset size 1..k): edges scanned */
  int64_t level_relax[8]; /*   and state-edge relaxations */
  int32_t grow_rounds;    /* grow rounds executed (one select + grow launch set each) */
  /* Progress (the anytime view of PrunedDP++, code/PrunedDP++/include/CPUNONHOP.h:3-5,
   * 684-726). gst_poll returns GST_PENDING with these filled for a query in flight;
   * a final result has ub == lb == cost when status is GST_OK. */
  int32_t ub;             /* weight of the best feasible tree found so far (-1: none yet);
                             never increases between polls and is >= the optimum */
  int32_t lb;             /* lower bound on the optimum (0 until the group distances are
                             final); never decreases and is <= the optimum */
  int32_t level;          /* DP level (subset size) the query has reached */
  int32_t pad_;
} gst_result;

typedef struct {
  int32_t n;
  int64_t m;
  int64_t graph_bytes;     /* resident CSR + groups */
  int64_t budget_bytes;    /* DP budget of the batched/served forms */
  int64_t in_use_bytes;    /* DP memory held by admitted queries */
  int32_t active_jobs, queued_jobs;
  int64_t completed;
} gst_info;

typedef struct gst_ctx gst_ctx; /* opaque; one per graph */

int gst_open(const gst_csr *csr, const gst_groups *groups, const gst_opts *opts, gst_ctx **ctx);
int gst_solve(gst_ctx *ctx, const int32_t *gids, int32_t k, int32_t D, int32_t flags,
              gst_result *out);
int gst_solve_batch(gst_ctx *ctx, const gst_query *queries, int32_t n, gst_result *out);
int gst_submit(gst_ctx *ctx, const gst_query *queries, int32_t n, uint64_t *tickets);
int gst_poll(gst_ctx *ctx, uint64_t ticket, gst_result *out);
int gst_wait(gst_ctx *ctx, uint64_t ticket, gst_result *out);
int gst_info_get(gst_ctx *ctx, gst_info *info);
/* Tunables of the grow phase (degree split in [1, 2^24], chunk size in [64, 2^24] arcs,
 * near-far step, 0 = off); used by the autotuner. Call while no batched or served query
 * is in flight. */
int gst_set_tuning(gst_ctx *ctx, int32_t degree_split, int32_t chunk_edges, int32_t near_far_delta);
/* The same per k (queries of k groups after de-duplication), plus the merge tile: at most
 * merge_tile vertices per merge work item (0 = as many as fit the LDS). Queries of other k
 * keep their tuning; a batch mixing k runs each job with its own. Used by the autotuner,
 * which keys its picks by (k, degree profile). Call while no batched or served query is in
 * flight. */
int gst_set_tuning_k(gst_ctx *ctx, int32_t k, int32_t degree_split, int32_t chunk_edges, int32_t merge_tile);
void gst_result_free(gst_result *r);
void gst_close(gst_ctx *ctx);
const char *gst_strerror(int code);
const char *gst_last_error(void);
int64_t gst_arena_bytes(int32_t n, int32_t k, int32_t D);
/* No counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). gst_solve without any
 * device or host allocation: the query's DP memory comes from a static device array of
 * GST_NOALLOC_MAX_BYTES per device (queries that need more return GST_OVER_BUDGET), and
 * the tree's distinct edges go to the caller's edges[3 * edge_cap] (out->edges stays
 * NULL; n - 1 edges always suffice; a larger tree returns GST_TREE_OMITTED with the cost).
 * Calls through one context are serialised; calls through two contexts on one device at
 * the same time race on the shared array. */
#define GST_NOALLOC_MAX_BYTES ((int64_t)64 << 20)
int gst_solve_noalloc(gst_ctx *ctx, const int32_t *gids, int32_t k, int32_t D, int32_t flags, int32_t *edges,
                      int32_t edge_cap, gst_result *out);
/* Device memory one admitted query of this shape holds in the batched and served forms
 * (arena, activation masks, worklists, tables, tree buffers): what admission charges
 * against dp_budget_bytes. < 0 on bad arguments. */
int64_t gst_job_bytes(gst_ctx *ctx, int32_t k, int32_t D);
/* No counterpart in the reference: instrumentation of the batched and served forms'
 * cohorts (queries of one k whose level 1 grows together in one shared level block).
 * Totals over every cohort since gst_open or the last reset (reset != 0 clears them after
 * the copy). */
typedef struct gst_cohort_stats {
  int64_t cohorts;        /* cohorts run */
  int64_t members;        /* queries they held */
  int64_t grow_launches;  /* level-1 grow launches (one per round) */
  double grow_device_s;   /* their device time (HIP events around each launch) */
  int64_t grow_arcs;      /* arcs they scanned (one scan serves every member) */
  int64_t grow_relax;     /* label relaxations they made */
} gst_cohort_stats;
int gst_get_cohort_stats(gst_ctx *ctx, gst_cohort_stats *out, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif
