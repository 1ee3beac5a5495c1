// gst_kernels.hpp — CDNA4 (gfx950, wave64) device code of the GST DP engine.
//
// Every kernel is multi-job: k_step advances each job (one query) to its phase for the
// round and lays out the round's work items flattened over jobs (Plan: per-kind prefix
// sums), so one launch with a fixed grid covers the work of every job and a block finds
// an item's job by binary search. The same kernels serve the single-query path (host
// reads the job state after every round, as the reference does at
// code/TrimCDP/src/DPQ.cu:512-544) and the batched path (rounds captured in a HIP graph,
// no host round trips).
//
// DP state T(v,p[,d]) = cost of the cheapest tree rooted at v that covers the query
// groups in subset p (and has height d, hop variant). Arena layout: int32, level-major:
// level L = |p| is one block holding, for every (vertex, height) row, the C(k,L) subsets
// of that size contiguously (ascending p, so bit j of a level mask is offset j):
//   T[off[L] + ((v*H + d) * C(k,L)) + rank(p)],  H = D+1 (1 without a diameter bound).
// A grow round of level L touches only that block (k=5: 20-40 B per vertex row), so a
// relaxation reads one 128-B line of the target and the level's working set is a
// fraction of the arena.
//
// Non-hop schedule (level = popcount of p):
//   level 1 : k-source SSSP = the singleton states (grow rounds)          [SSSP, a4]
//   tables  : group distance table and Held-Karp tour tables              [a5]
//   level L : merge kernel (row staged in LDS), then grow rounds          [a8]
//   level k : merge only; answer = min_v T(v, full)                       [a9]
// Pruning (all admissible, so the optimum is preserved):
//   * upper bound UB = min over realizable trees: T(v,p) + sum_{i not in p} T(v,{i})
//     (tree plus shortest paths) and T(v,p) + T(v,~p); the reference uses the same
//     two updates (DPQ.cu:82-107);
//   * a state is grown only if cost <= UB/2: in an optimal tree rooted at its weighted
//     centroid every grown subtree weighs at most half the optimum (the reference's
//     "cost-1 > best/2" cut, DPQ.cu:108, and CPUNONHOP.h:795);
//   * a state is grown only if cost + lb <= UB, lb = max(one-label, tour, tour-2)
//     bounds computed from lower bounds of the group distances (DPQ.cu:174-224,
//     CPUNONHOP.h:436-496).
//   Level-1 growth uses only the UB/2 rule, so every singleton value <= UB_final/2 is an
//   exact distance and dl_i(v) = min(T(v,{i}), UB_final/2 + 1) is a valid lower bound.
// Hop schedule: per level the merge kernel fills all heights, then heights 0..D-1 are
//   grown in order (a grow raises the height by exactly one, so each height is final
//   before it is grown). A stored hop state holds a tree of height at most d. Pruning
//   (the reference's D-PrunedDP++ and D-TrimCDP prune nothing): cost + lb > UB with lb
//   from D-d hop group distances and UB from stars and complement merges
//   (grow_select_hop); no UB/2 rule, since re-rooting changes heights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gst {

constexpr int32_t INF = 0x3fffffff;   // unreachable; sum of two INF fits int32
constexpr int32_t LB_CAP = 1 << 24;   // lower bounds are clamped here (still admissible)
constexpr int BLK = 256;              // threads per block (4 wavefronts)
constexpr int WAVE = 64;
constexpr int MAXK = 7;
constexpr int MAXW = 1 << MAXK;
constexpr int MAXLISTS = 16;          // non-hop: 2 (parity); hop: D+1 <= 16
constexpr int MAXJOBS = 256;          // jobs one launch can carry
constexpr int FINAL_CHUNK = 1024;     // arena rows per k_final work item

enum Phase : int32_t {
  PH_INIT = 0, PH_SEED = 1, PH_GROW = 2, PH_TABLES = 3, PH_MERGE = 4,
  PH_FINAL = 5, PH_REBUILD = 6, PH_HOPCLEAR = 7, PH_DONE = 8, PH_FARSPLIT = 9
};

enum Stat : int32_t {
  ST_VERTS = 0,   // frontier vertices processed (|S|)
  ST_EDGES = 1,   // edges scanned (E_S)
  ST_STATES = 2,  // frontier states evaluated (|F|)
  ST_RELAX = 3,   // state-edge relaxations (R)
  ST_PUSH = 4,    // worklist pushes (|F_next|, vertices)
  ST_ATOM = 5,    // atomicMin issued on the arena
  ST_GROWR = 6,   // grow rounds
  ST_MERGE = 7    // merged states written
};
// Work counters are striped: a block adds to stripe blockIdx % NSTRIPE, so the per-wave
// flushes of one round spread over NSTRIPE cache lines instead of queueing on one word.
constexpr int NSTRIPE = 64;
constexpr int NSTX = 32;       // counters per stripe: Stat values, then per-level pairs
constexpr int SX_LEVEL = 8;    // level L: edges at SX_LEVEL + 2L, relaxations at SX_LEVEL + 2L + 1

struct JobConst {
  int32_t *T;                 // arena
  unsigned long long *stx;    // [NSTRIPE][NSTX] work counters, summed by the host
  unsigned long long *act;    // [nl][n] active-subset masks (bit j = j-th subset of the level)
  int32_t *lst;               // [nl][n] thread-per-vertex worklists
  int32_t *blst;              // [nl][bcap][2] (vertex, chunk) wavefront work items
  int32_t *tab;               // tour tables: hk[K][K][W], w2[K][W], vv[K][K]
  int32_t *stk;               // rebuild stack: scap x (v, p, d)
  int32_t *out;               // rebuilt tree: ecap x (u, v, w)
  long long bcap;
  int32_t scap, ecap;
  int32_t K, D, H, W, R, hop, flags, nl;
  int32_t seed_total;         // members of the query groups (seed work items)
  int32_t tv;                 // vertices per merge tile
  int32_t tvb;                // vertices per bound-first hop merge tile (a multiple of tv)
  int32_t delta;              // near-far step of the grow frontier (0 = off)
  int32_t passive;            // cohort member: the grow phases run in the cohort kernels
  int32_t sparse;             // merges of levels >= 2 cover only the vertices of alv (below)
  int32_t *alv;               // [n] vertices that can lie on an optimal tree (k_alive): the
                              // arena's unused tail (levels 1..k take n * (W - 1) words)
  unsigned long long *alb;    // [ceil(n / 64)] bit v: v is listed in alv; grows of levels >= 2
                              // skip targets whose bit is clear
  // grow tuning of this job (chosen per k): vertices of degree >= split are grown as
  // ceil(deg / chunk) wavefront items (nchunk[v]; 0 = thread item), the others by threads;
  // with vertices numbered by descending degree those are the ids below n_heavy
  const int32_t *nchunk;
  int32_t split, chunk, n_heavy;
  // cohort member only: the merge of a level L < K writes its states into the cohort's
  // level block cT (row of vertex v at v * S, S = cohort_stride(cB, C(k,L)), this member's
  // labels at cq * C(k,L)) and activates them in list 1 of the cohort's label masks
  // (cmask: [2][n][S / 64]) and vertex flags (cflag: [2][n])
  int32_t *cT;
  unsigned long long *cmask;
  int32_t *cflag;
  int32_t cB, cq;             // members in the cohort, this member's index
  long long off[9];           // arena block offset of each level (subset size)
  int32_t gid[8];
  uint8_t nsub[8];
  uint8_t sub[8][64];         // sub[L][j]: j-th subset of popcount L, ascending
  uint8_t rank[128];          // rank of p within its popcount level
};

struct alignas(16) JobState {
  int32_t phase, level, depth, parity;
  int32_t cnt[MAXLISTS][2];   // [list][0 = thread items, 1 = wavefront items]
  int32_t ub, r1, err, n_out;
  unsigned long long ans;     // (cost << 32) | (v*H + d), min-reduced
  int32_t rounds, dp_rounds;
  int32_t theta;              // near-far threshold: states above it wait in the far pile
  int32_t far_min;            // min cost deferred to the far pile this round
  int32_t fpar;               // far pile parity (lists 2 + fpar)
  int32_t minc;               // min cost activated by the last merge (level start)
  int32_t sat;                // a state came within reach of INF (costs may have saturated)
  int32_t lb;                 // lower bound on the optimum (set after level 1; 0 before)
  int32_t n_alive;            // entries of alv (-1: not built, the merges cover every vertex)
  unsigned long long st[8];
  unsigned long long stl[8][2];  // per level: edges scanned, state-edge relaxations
};

// Non-hop list indices: near frontier 0/1 (parity), far pile 2/3 (fpar); mask arrays:
// near 0/1, far pile 2.
constexpr int FAR_LIST = 2;
constexpr int FAR_ACT = 2;

// Work items of one round, flattened over jobs: off[kind][j] = first item of job j.
// WK_MERGEB: the bound-first tiles of hop merges (k_merge_hop<true>); WK_MERGE: the others.
// WK_ALIVE: the vertices k_alive tests after the tables of a job with sparse merges.
enum WorkKind : int32_t { WK_SEED = 0, WK_THR = 1, WK_WAVE = 2, WK_CLEAR = 3, WK_MERGE = 4, WK_FINAL = 5, WK_FAR = 6,
                          WK_COLLECT = 7, WK_MERGEB = 8, WK_ALIVE = 9, NWK = 10 };
struct Plan {
  int32_t njobs, pad;
  long long off[NWK][MAXJOBS + 1];
};

struct GraphDev {
  const int32_t *beg;   // [n+1]
  const int32_t *adj;   // [m]
  const int32_t *w;     // [m]
  const long long *goffs;
  const int32_t *gverts;
  int32_t n;
};

// Costs saturate at INF: a grow step c + w or a merge c1 + c2 at or above INF is stored
// as INF (unreachable). A finite answer is always exact (every state of an optimal
// derivation costs at most the optimum). A saturated sum needs a finite state of cost
// >= SAT_AT (a merge half, or a grow source, since weights are <= 2^20 < SAT_AT), and
// every finite state passes the select pass (when grown) or a merge (when merged), which
// flag the job then, so that a "no solution" outcome is reported as a cost-range failure.
constexpr int32_t SAT_AT = INF >> 1;
__device__ __forceinline__ void flag_saturation(JobState *s) {
  if (!__hip_atomic_load(&s->sat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    __hip_atomic_fetch_or(&s->sat, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// nchunk[u] without a memory access for light vertices: with vertices numbered by
// descending degree, those of degree >= split are exactly the ids below n_heavy.
__device__ __forceinline__ int32_t chunks_of(const int32_t *nchunk, int32_t n_heavy, int32_t u) {
  return u < n_heavy ? nchunk[u] : 0;
}
__device__ __forceinline__ int32_t chunks_of(const JobConst &c, int32_t u) {
  return chunks_of(c.nchunk, c.n_heavy, u);
}

constexpr int JF_NO_PRUNE = 4;  // == GST_F_NO_PRUNE
constexpr int JF_NO_TREE = 2;   // == GST_F_NO_TREE

__device__ __forceinline__ int popc(uint32_t x) { return __builtin_popcount(x); }

// Device memory through a global-address-space pointer: an access through a generic
// pointer is a flat access, which also counts against the LDS/scalar counter, so every
// later LDS or scalar wait would wait for it as well.
typedef __attribute__((address_space(1))) int32_t gi32;
typedef __attribute__((address_space(1))) unsigned long long gu64;
__device__ __forceinline__ gi32 *as_global(int32_t *p) { return (gi32 *)p; }
__device__ __forceinline__ const gi32 *as_global(const int32_t *p) { return (const gi32 *)p; }
__device__ __forceinline__ gu64 *as_global(unsigned long long *p) { return (gu64 *)p; }
__device__ __forceinline__ const gu64 *as_global(const unsigned long long *p) { return (const gu64 *)p; }

// Words of a cohort row: B members' C(k,L) labels, padded to whole 64-label chunks.
__host__ __device__ __forceinline__ int cohort_stride(int B, int C) { return ((B * C + 63) / 64) * 64; }

// Arena index of state (v, p, d).
__device__ __forceinline__ size_t sidx(const JobConst &c, int32_t v, int p, int d) {
  const int L = popc(p);
  return (size_t)c.off[L] + ((size_t)v * c.H + d) * c.nsub[L] + c.rank[p];
}
// Row of (v, d) in level L's block: C(k,L) contiguous states.
__device__ __forceinline__ int32_t *lrow(const JobConst &c, int L, int32_t v, int d) {
  return c.T + c.off[L] + ((size_t)v * c.H + d) * c.nsub[L];
}

__device__ __forceinline__ int load_ub(const JobState *s) {
  return __hip_atomic_load(&s->ub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

#ifndef GST_L1_FULL
#define GST_L1_FULL 1
#endif
// Largest cost at which a non-hop state of level L is grown: floor(UB/2), the centroid
// rule, at levels >= 2; level 1 grows up to UB - 1, so that its distances are exact up to
// UB (a shortest path's last source costs at most its length - 1) and the lower bounds
// built on them (lb_tail, the tour tables) are clamped at UB + 1 instead of UB/2 + 1.
// Growing more at level 1 is always safe (it only makes more values exact); on the C4
// workload it costs 15% more level-1 relaxations and removes 90% of those of levels 2-4.
// GST_L1_FULL=0 restores the centroid rule at level 1 as well.
__device__ __forceinline__ int grow_cap(int ub, int L) {
  if (GST_L1_FULL && L == 1) return ub >= INF ? ub : ub - 1;
  return ub >> 1;
}

// list index of the worklist a job is reading (cur) / writing (nxt) this round
__device__ __forceinline__ int cur_list(const JobConst &c, const JobState *s) {
  return c.hop ? s->depth : s->parity;
}
__device__ __forceinline__ int nxt_list(const JobConst &c, const JobState *s) {
  return c.hop ? s->depth + 1 : (s->parity ^ 1);
}

// Append vertex u to worklist li: a thread item when nch == 0, otherwise nch wavefront
// items (one per chunk of edges). The far piles of the near-far frontier take thread
// items only, whatever the degree, so callers pass nch = 0 for them.
__device__ __forceinline__ void push_item(const JobConst &c, JobState *s, const GraphDev &g,
                                          int li, int32_t u, int nch) {
  if (nch == 0) {
    const int pos = atomicAdd(&s->cnt[li][0], 1);
    c.lst[(size_t)li * g.n + pos] = u;
  } else {
    const int pos = atomicAdd(&s->cnt[li][1], nch);
    int32_t *b = c.blst + ((size_t)li * c.bcap + pos) * 2;
    for (int i = 0; i < nch; i++) { b[2 * i] = u; b[2 * i + 1] = i; }
  }
}

// Append vertex u to worklist li by its degree class: a thread item when deg < split,
// otherwise one wavefront item per chunk of edges.
__device__ __forceinline__ void push_vertex(const JobConst &c, JobState *s, const GraphDev &g,
                                            int li, int32_t u) {
  const int32_t d = g.beg[u + 1] - g.beg[u];
  push_item(c, s, g, li, u, d < c.split ? 0 : (d + c.chunk - 1) / c.chunk);
}

// Admissible lower bound on the weight still needed to connect a tree rooted at v
// (covering p) to the groups outside p. sing[i] = T(v, {i}) (group distances).
__device__ int lb_tail(const int32_t *sing, int32_t r1, int p, const JobConst &c,
                       const int32_t *hk, const int32_t *w2) {
  const int X = c.R ^ p;
  if (X == 0) return 0;
  const int K = c.K, W = c.W;
  int dl[MAXK];
#pragma unroll
  for (int i = 0; i < MAXK; i++) {
    int x = (i < K) ? sing[i] : INF;
    x = x < r1 ? x : r1;
    dl[i] = x < LB_CAP ? x : LB_CAP;
  }
  int one = 0, mind = INF, lb1 = INF, lb2 = 0;
#pragma unroll
  for (int a = 0; a < MAXK; a++) {
    if (a < K && ((X >> a) & 1)) {
      one = dl[a] > one ? dl[a] : one;
      mind = dl[a] < mind ? dl[a] : mind;
    }
  }
#pragma unroll
  for (int a = 0; a < MAXK; a++) {
    if (!(a < K && ((X >> a) & 1))) continue;
#pragma unroll
    for (int b = 0; b < MAXK; b++) {
      if (!(b < K && ((X >> b) & 1))) continue;
      const int t = (dl[a] + hk[(a * K + b) * W + X] + dl[b]) >> 1;
      lb1 = t < lb1 ? t : lb1;
    }
    const int t2 = (dl[a] + w2[a * W + X] + mind) >> 1;
    lb2 = t2 > lb2 ? t2 : lb2;
  }
  int lb = lb1 > lb2 ? lb1 : lb2;
  return lb > one ? lb : one;
}

// Realizable upper bounds from state (v,p) with cost cst: tree + shortest paths to the
// missing groups, and tree + complementary tree when that level is already final.
__device__ __forceinline__ int ub_candidate(const int32_t *sing, int p, int cst, int L,
                                            const JobConst &c, int compv) {
  long long star = cst;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < MAXK; i++)
    if (i < c.K && !((p >> i) & 1)) {
      const int x = sing[i];
      ok = ok && x < INF;
      star += x;
    }
  long long best = ok ? star : (long long)INF;
  if (compv < INF && (long long)cst + compv < best) best = (long long)cst + compv;
  return best < INF ? (int)best : INF;
}

// --------------------------------------------------------------------------- step
// Turn the result of the previous round into this job's phase for this round.
__device__ void step_job(const JobConst &c, JobState &s) {
  int ph = s.phase;
  if (ph == PH_DONE) return;
  s.rounds++;
  if (c.passive) return;  // a cohort member: its driver sets the phase of every round
  switch (ph) {
    case PH_INIT:
      s.level = 1; s.depth = 0; s.parity = 0;
      s.phase = PH_SEED;
      return;
    case PH_SEED:
      if (c.K == 1) { s.phase = PH_FINAL; return; }
      if (c.hop) { s.depth = 0; goto hop_entry; }
      goto level_start;
    case PH_GROW:
      if (c.hop) { s.depth++; goto hop_entry; }
      goto grow_entry;
    case PH_TABLES:
      s.level = 2;
      s.phase = PH_MERGE;
      return;
    case PH_MERGE:
      if (s.level == c.K) { s.phase = PH_FINAL; return; }
      if (c.hop) { s.depth = 0; goto hop_entry; }
      goto level_start;
    case PH_FARSPLIT:
      s.cnt[FAR_LIST + s.fpar][0] = 0;  // consumed; the re-deferred rest is in the other pile
      s.fpar ^= 1;
      goto grow_entry;
    case PH_HOPCLEAR:
      for (int i = 0; i < MAXLISTS; i++) { s.cnt[i][0] = 0; s.cnt[i][1] = 0; }
      if (s.level == 1) { s.phase = PH_TABLES; return; }  // lower bound from the distances
      s.level++;
      s.depth = 0;
      s.phase = PH_MERGE;
      return;
    case PH_FINAL:
      s.phase = (s.ans != ~0ull && !(c.flags & JF_NO_TREE)) ? PH_REBUILD : PH_DONE;
      return;
    case PH_REBUILD:
      s.phase = PH_DONE;
      return;
    default:
      s.err = 99;
      s.phase = PH_DONE;
      return;
  }
level_start:
  s.theta = c.delta > 0 ? (s.minc < INF ? s.minc : 0) + c.delta : INF;
  s.minc = INF;
  s.far_min = INF;
grow_entry: {
    const int nx = s.parity ^ 1;
    if (s.cnt[nx][0] + s.cnt[nx][1] == 0 && s.cnt[FAR_LIST + s.fpar][0] > 0) {
      // near frontier drained: raise the threshold past the cheapest waiting state
      s.cnt[s.parity][0] = 0; s.cnt[s.parity][1] = 0;
      s.theta = (s.far_min > s.theta ? s.far_min : s.theta) + c.delta;
      s.far_min = INF;
      s.phase = PH_FARSPLIT;
      return;
    }
    if (s.cnt[nx][0] + s.cnt[nx][1] == 0) {  // level converged
      s.cnt[s.parity][0] = 0; s.cnt[s.parity][1] = 0;
      if (s.level == 1) { s.phase = PH_TABLES; return; }
      s.level++;
      s.phase = PH_MERGE;
      return;
    }
    s.cnt[s.parity][0] = 0; s.cnt[s.parity][1] = 0;  // consumed last round
    s.parity = nx;
    s.phase = PH_GROW;
    s.dp_rounds++;
    return;
  }
hop_entry: {
    while (s.depth < c.D && s.cnt[s.depth][0] + s.cnt[s.depth][1] == 0) s.depth++;
    if (s.depth >= c.D) { s.phase = PH_HOPCLEAR; return; }
    s.phase = PH_GROW;
    s.dp_rounds++;
    return;
  }
}

// Hop merge tiles of levels 4..k-1 (8 or more splits per subset) take the bound-first form
// (k_merge_hop<true>, work kind WK_MERGEB) when pruning is on, over tvb vertices per tile.
__host__ __device__ __forceinline__ bool hop_bound_first(const JobConst &c, int L) {
  return c.hop && !(c.flags & JF_NO_PRUNE) && L >= 4 && L < c.K;
}

// One block (BLK >= njobs threads): advance every job, then lay out this round's work
// items per kind as exclusive prefix sums over jobs.
__global__ __launch_bounds__(BLK) void k_step(const JobConst *__restrict__ jcs, JobState *jss, Plan *pl,
                                              int njobs, int n) {
  __shared__ long long cnt[NWK][MAXJOBS];
  const int j = threadIdx.x;
  if (j < njobs) {
    const JobConst &c = jcs[j];
    JobState &s = jss[j];
    step_job(c, s);
    const int ph = s.phase;
    long long k[NWK] = {};
    if (ph == PH_SEED) k[WK_SEED] = c.seed_total;
    if (ph == PH_GROW) {
      const int li = c.hop ? s.depth : s.parity;
      k[WK_THR] = s.cnt[li][0];
      k[WK_WAVE] = s.cnt[li][1];
      k[WK_CLEAR] = k[WK_THR] + k[WK_WAVE];
    }
    if (ph == PH_HOPCLEAR) k[WK_CLEAR] = s.cnt[c.D][0] + s.cnt[c.D][1];
    // padded to whole wavefronts: every wavefront of k_alive then tests 64 vertices of one
    // job, which are one word of its alive bitmap
    if (ph == PH_TABLES && c.sparse && !c.passive) k[WK_ALIVE] = ((long long)n + WAVE - 1) & ~(long long)(WAVE - 1);
    if (ph == PH_MERGE) {
      if (hop_bound_first(c, s.level)) k[WK_MERGEB] = ((long long)n + c.tvb - 1) / c.tvb;
      else {
        const long long nv = (c.sparse && !c.passive && s.n_alive >= 0) ? s.n_alive : n;
        k[WK_MERGE] = (nv + c.tv - 1) / c.tv;
      }
    }
    // blocks of 64 vertices per list the merge activates into (heights 0..D-1, or one list)
    if (ph == PH_MERGE && s.level < c.K) k[WK_COLLECT] = (((long long)n + 63) / 64) * (c.hop ? c.D : 1);
    if (ph == PH_FINAL) k[WK_FINAL] = ((long long)n * c.H + FINAL_CHUNK - 1) / FINAL_CHUNK;
    if (ph == PH_FARSPLIT) k[WK_FAR] = s.cnt[FAR_LIST + s.fpar][0];
#pragma unroll
    for (int q = 0; q < NWK; q++) cnt[q][j] = k[q];
  }
  __syncthreads();
  if (j < NWK) {
    long long acc = 0;
    pl->off[j][0] = 0;
    for (int t = 0; t < njobs; t++) { acc += cnt[j][t]; pl->off[j][t + 1] = acc; }
  }
  if (j == 0) pl->njobs = njobs;
}

// Copy the offsets of one work kind into LDS; returns the number of jobs.
__device__ __forceinline__ int load_plan(const Plan *pl, int kind, long long *soff) {
  const int nj = pl->njobs;
  for (int t = threadIdx.x; t <= nj; t += blockDim.x) soff[t] = pl->off[kind][t];
  __syncthreads();
  return nj;
}
// Job owning flattened item i: the largest j with soff[j] <= i.
__device__ __forceinline__ int find_job(const long long *soff, int nj, long long i) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (soff[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// --------------------------------------------------------------------------- fill
__global__ void k_fill_i32(int32_t *p, long long n, int32_t v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long n4 = n >> 2;
  int4 *p4 = reinterpret_cast<int4 *>(p);
  const int4 v4 = make_int4(v, v, v, v);
  for (long long k = i; k < n4; k += stride) p4[k] = v4;
  for (long long k = (n4 << 2) + i; k < n; k += stride) p[k] = v;
}

__global__ void k_fill_u64(unsigned long long *p, long long n, unsigned long long v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += stride) p[k] = v;
}

// --------------------------------------------------------------------------- seed
// T(x, {i}) = 0 for every member x of query group i (height 0); activates bit i.
__global__ __launch_bounds__(BLK) void k_seed(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  const int nj = load_plan(pl, WK_SEED, soff);
  const long long total = soff[nj];
  for (long long it = (long long)blockIdx.x * BLK + threadIdx.x; it < total; it += (long long)gridDim.x * BLK) {
    const int jb = find_job(soff, nj, it);
    const JobConst &c = jcs[jb];
    JobState *s = &jss[jb];
    long long loc = it - soff[jb];
    int i = 0;
    for (; i < c.K; i++) {
      const long long sz = g.goffs[c.gid[i] + 1] - g.goffs[c.gid[i]];
      if (loc < sz) break;
      loc -= sz;
    }
    if (i >= c.K) continue;
    const int li = c.hop ? 0 : 1;
    const int32_t x = g.gverts[g.goffs[c.gid[i]] + loc];
    lrow(c, 1, x, 0)[i] = 0;
    const unsigned long long old = atomicOr(&c.act[(size_t)li * g.n + x], 1ull << i);
    if (old == 0) push_vertex(c, s, g, li, x);
  }
}

// --------------------------------------------------------------------------- grow
// Evaluate the active states of v (bits of a) at the current level: update UB, and
// return the subset of bits that may grow. Non-hop jobs (grow_select_hop for hop jobs).
// Non-hop: v's active states of level L are bits of a (bit j = offset j of v's level row).
// row: v's states of level L (bit j of a = row[j]); sing: v's singleton states.
__device__ __forceinline__ unsigned long long grow_select_rows(const JobConst &c, JobState *s, int32_t v,
                                                               const int32_t *row, const int32_t *sing,
                                                               unsigned long long a, int L, int ub, int &n_states,
                                                               int theta, unsigned long long &defer, int &dmin) {
  const int32_t *hk = c.tab;
  const int32_t *w2 = c.tab + c.K * c.K * c.W;
  const bool prune = !(c.flags & JF_NO_PRUNE);
  // the group distances and r1 are read once per vertex, not once per state
  int sl[MAXK];
#pragma unroll
  for (int i = 0; i < MAXK; i++) sl[i] = i < c.K ? sing[i] : INF;
  const int32_t r1 = s->r1;
  unsigned long long gm = 0;
  int ubl = ub;
  while (a) {
    const int j = __builtin_ctzll(a);
    a &= a - 1;
    n_states++;
    const int p = c.sub[L][j];
    const int cst = row[j];
    if (cst >= INF) continue;
    if (cst >= SAT_AT) flag_saturation(s);
    const int q = c.R ^ p;
    const int compv = (q && popc(q) < L) ? c.T[sidx(c, v, q, 0)] : INF;
    const int cand = ub_candidate(sl, p, cst, L, c, compv);
    if (cand < ubl) {
      atomicMin(&s->ub, cand);
      ubl = cand;
    }
    if (prune) {
      if (cst > grow_cap(ubl, L)) continue;
      if (L > 1 && cst + lb_tail(sl, r1, p, c, hk, w2) > ubl) continue;
    }
    if (cst > theta) {  // near-far: wait in the far pile
      defer |= 1ull << j;
      dmin = cst < dmin ? cst : dmin;
      continue;
    }
    gm |= 1ull << j;
  }
  return gm;
}

__device__ __forceinline__ unsigned long long grow_select(const JobConst &c, JobState *s,
                                                          int32_t v, unsigned long long a,
                                                          int L, int ub, int &n_states, int theta,
                                                          unsigned long long &defer, int &dmin) {
  return grow_select_rows(c, s, v, lrow(c, L, v, 0), lrow(c, 1, v, 0), a, L, ub, n_states, theta, defer, dmin);
}

// Diameter-bounded variant: evaluate the active states of (v, level L, height d).
// UB candidates (each is the value of a DP merge, so UB stays >= the DP optimum):
// - the star of v: merging the singletons (v, {i}, h_i) with every h_i <= floor(D/2) is
//   legal in any order (any two heights sum to <= D), so sum_i min_{h <= D/2} T(v,{i},h);
// - a state with a final complement q = R \ p: T(v,p,d) + min_{d2 <= D-d} T(v,q,d2).
// Lower bound (levels >= 2): on an optimal derivation through (v,p,d), the rest of the
// derivation holds, for each group i outside p, a walk from v to a member of i of at
// most D - d hops (up from v to the merge point, then down a branch; their heights add
// up to <= D), so its cost is >= max_{i not in p} min_{h <= D-d} T(v,{i},h). Level 1 is
// not pruned, so these distances are exact when levels >= 2 run. No UB/2 rule here: the
// centroid re-rooting behind it changes heights.
__device__ unsigned long long grow_select_hop(const JobConst &c, JobState *s, int32_t v,
                                              unsigned long long a, int L, int d, int ub,
                                              int &n_states) {
  const int K = c.K, D = c.D, H = c.H;
  const int32_t *l1 = lrow(c, 1, v, 0);  // H rows of K group distances, contiguous
  int dlb[MAXK], dst[MAXK];
  // pm: prefix minima over the heights read so far; lo[i][x] = min_{h <= x} T(v,{i},h)
  // for x <= HX, kept for the long-arm stars below
  constexpr int HX = 2;
  int pm[MAXK], lo[MAXK][HX + 1];
  long long arm = INF;  // best long-arm star of v
#pragma unroll
  for (int i = 0; i < MAXK; i++) {
    dlb[i] = INF; dst[i] = INF; pm[i] = INF;
#pragma unroll
    for (int x = 0; x <= HX; x++) lo[i][x] = INF;
  }
  for (int h0 = 0; h0 < H; h0 += 4) {  // 4 heights of loads in flight at a time
    int xs[4][MAXK];
#pragma unroll
    for (int hh = 0; hh < 4; hh++)
#pragma unroll
      for (int i = 0; i < MAXK; i++) xs[hh][i] = (h0 + hh < H && i < K) ? l1[(h0 + hh) * K + i] : INF;
#pragma unroll
    for (int hh = 0; hh < 4; hh++) {
      const int h = h0 + hh;
#pragma unroll
      for (int i = 0; i < MAXK; i++) {
        const int x = xs[hh][i];
        if (h <= D - d && x < dlb[i]) dlb[i] = x;
        if (2 * h <= D && x < dst[i]) dst[i] = x;
        if (h < H && x < pm[i]) pm[i] = x;
#pragma unroll
        for (int q = 0; q <= HX; q++)
          if (h == q) lo[i][q] = pm[i];
      }
      // Long-arm stars: one group j reached within a hops (a >= D/2), every other within
      // D - a. Merging these singletons at v in any order is legal in the DP (any two
      // heights sum to <= D), so each value is a DP state's value and bounds the optimum
      // from above; a = ceil(D/2) already beats the plain star for odd D.
      const int xo = D - h;
      if (h < H && 2 * h >= D && xo <= HX) {
        long long sum = 0;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < MAXK; i++) {
          if (i >= K) continue;
          int lv = INF;
#pragma unroll
          for (int q = 0; q <= HX; q++)
            if (xo == q) lv = lo[i][q];
          ok = ok && lv < INF;
          sum += lv;
        }
        if (ok) {
#pragma unroll
          for (int j = 0; j < MAXK; j++) {
            if (j >= K) continue;
            int lj = INF;
#pragma unroll
            for (int q = 0; q <= HX; q++)
              if (xo == q) lj = lo[j][q];
            const long long cand = sum - lj + pm[j];
            arm = cand < arm ? cand : arm;
          }
        }
      }
    }
  }
  int ubl = ub;
  if (arm < ubl) { atomicMin(&s->ub, (int)arm); ubl = (int)arm; }
  {
    long long star = 0;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < MAXK; i++)
      if (i < K) { ok = ok && dst[i] < INF; star += dst[i]; }
    if (ok && star < ubl) { atomicMin(&s->ub, (int)star); ubl = (int)star; }
  }
  const bool prune = !(c.flags & JF_NO_PRUNE) && L > 1;
  const int32_t *row = lrow(c, L, v, d);
  unsigned long long gm = 0;
  while (a) {
    const int j = __builtin_ctzll(a);
    a &= a - 1;
    n_states++;
    const int p = c.sub[L][j];
    const int cst = row[j];
    if (cst >= INF) continue;
    if (cst >= SAT_AT) flag_saturation(s);
    const int q = c.R ^ p;
    if (q && popc(q) < L) {
      const int Lq = popc(q);
      const int32_t *cq = lrow(c, Lq, v, 0) + c.rank[q];
      const int nq = c.nsub[Lq];
      int mq = INF;
      for (int d0 = 0; d0 <= D - d; d0 += 4) {  // 4 heights of loads in flight at a time
        int xs[4];
#pragma unroll
        for (int dd = 0; dd < 4; dd++) xs[dd] = d0 + dd <= D - d ? cq[(size_t)(d0 + dd) * nq] : INF;
#pragma unroll
        for (int dd = 0; dd < 4; dd++) mq = xs[dd] < mq ? xs[dd] : mq;
      }
      if (mq < INF && cst + mq < ubl) { atomicMin(&s->ub, cst + mq); ubl = cst + mq; }
    }
    if (prune) {
      int lb = 0;
#pragma unroll
      for (int i = 0; i < MAXK; i++)
        if (i < K && !((p >> i) & 1)) lb = dlb[i] > lb ? dlb[i] : lb;
      if ((long long)cst + lb > ubl) continue;
    }
    gm |= 1ull << j;
  }
  return gm;
}

// Up to RG active subsets of one source vertex are relaxed together over each edge: the
// RG target loads of an edge are issued back to back, an improvement is detected
// against the loaded value, and the atomicMin is fire-and-forget (its result is unused,
// so no round trip). A concurrent update that got lower first only makes the activation
// redundant: activating whenever cost < loaded value covers every real improvement.
constexpr int RG = 8;
struct RelaxGroup {
  int n;
  int p[RG];   // offsets in the level row (= bit of the level mask)
  int c[RG];   // source costs
};

// src = the source vertex's level row
__device__ __forceinline__ void take_group(const int32_t *srcp, unsigned long long &gm, RelaxGroup &G) {
  const gi32 *src = as_global(srcp);
  G.n = 0;
#pragma unroll
  for (int i = 0; i < RG; i++) {
    G.p[i] = 0; G.c[i] = INF;
    if (gm) {
      const int j = __builtin_ctzll(gm);
      gm &= gm - 1;
      G.p[i] = j;
      G.c[i] = src[j];
      G.n = i + 1;
    }
  }
}

// Make a group held identically by every lane live in scalar registers.
__device__ __forceinline__ void uniform_group(RelaxGroup &G) {
  G.n = __builtin_amdgcn_readfirstlane(G.n);
#pragma unroll
  for (int i = 0; i < RG; i++) {
    G.p[i] = __builtin_amdgcn_readfirstlane(G.p[i]);
    G.c[i] = __builtin_amdgcn_readfirstlane(G.c[i]);
  }
}

struct GrowCounters {
  int verts, edges, states, relax, push, atom;
  int dmin;  // cheapest state deferred to the far pile
  __device__ void zero() { verts = edges = states = relax = push = atom = 0; dmin = INF; }
};

// This block's stripe of counter i of a job.
__device__ __forceinline__ unsigned long long *stat_at(const JobConst &c, int i) {
  return c.stx + (blockIdx.x & (NSTRIPE - 1)) * NSTX + i;
}

// Add a thread's counters to job cj's stats: one atomic per counter per wavefront when
// all lanes worked for the same job (the common case), per lane otherwise.
__device__ __forceinline__ void flush_lane(const JobConst *jcs, JobState *jss, int cj, const GrowCounters &k) {
  if (cj < 0) return;
  const JobConst &c = jcs[cj];
  JobState *s = &jss[cj];
  const int L = s->level & 7;
  if (k.edges) atomicAdd(stat_at(c, SX_LEVEL + 2 * L), (unsigned long long)k.edges);
  if (k.relax) atomicAdd(stat_at(c, SX_LEVEL + 2 * L + 1), (unsigned long long)k.relax);
  if (k.verts) atomicAdd(stat_at(c, ST_VERTS), (unsigned long long)k.verts);
  if (k.edges) atomicAdd(stat_at(c, ST_EDGES), (unsigned long long)k.edges);
  if (k.states) atomicAdd(stat_at(c, ST_STATES), (unsigned long long)k.states);
  if (k.relax) atomicAdd(stat_at(c, ST_RELAX), (unsigned long long)k.relax);
  if (k.push) atomicAdd(stat_at(c, ST_PUSH), (unsigned long long)k.push);
  if (k.atom) atomicAdd(stat_at(c, ST_ATOM), (unsigned long long)k.atom);
  if (k.dmin < INF) atomicMin(&s->far_min, k.dmin);
}
// Called by every lane of the wavefront (converged).
__device__ void flush_counters(const JobConst *jcs, JobState *jss, int cj, GrowCounters &k) {
  const int first = __shfl(cj, 0, 64);
  const bool same = __all(cj == first || cj < 0);
  if (!same || first < 0) { flush_lane(jcs, jss, cj, k); return; }
  int dm = k.dmin;
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(dm, off, 64);
    dm = o < dm ? o : dm;
  }
  unsigned long long v[6] = {(unsigned long long)k.verts, (unsigned long long)k.edges,
                             (unsigned long long)k.states, (unsigned long long)k.relax,
                             (unsigned long long)k.push, (unsigned long long)k.atom};
#pragma unroll
  for (int q = 0; q < 6; q++) {
    unsigned long long x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[q] = x;
  }
  if ((threadIdx.x & 63) == 0) {
    const JobConst &c = jcs[first];
    JobState *s = &jss[first];
    const int L = s->level & 7;
    if (v[1]) atomicAdd(stat_at(c, SX_LEVEL + 2 * L), v[1]);
    if (v[3]) atomicAdd(stat_at(c, SX_LEVEL + 2 * L + 1), v[3]);
    const int map[6] = {ST_VERTS, ST_EDGES, ST_STATES, ST_RELAX, ST_PUSH, ST_ATOM};
#pragma unroll
    for (int q = 0; q < 6; q++)
      if (v[q]) atomicAdd(stat_at(c, map[q]), v[q]);
    if (dm < INF) atomicMin(&s->far_min, dm);
  }
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Wavefront-aggregated append of cnt items (cnt = 0: nothing) to the counter *ctr that
// every active lane shares; one atomicAdd per wavefront. Must be called converged.
__device__ __forceinline__ int wave_append(int *ctr, int cnt) {
  const int lane = lane_id();
  int x = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  const int total = __shfl(x, 63, 64);
  int base = 0;
  if (lane == 63 && total) base = atomicAdd(ctr, total);
  base = __shfl(base, 63, 64);
  return base + x - cnt;
}

// Per-wavefront worklist staging in LDS: pushes are buffered and appended to the global
// list with one atomicAdd per flush instead of one per step. Every wavefront of every
// job hits the same two list counters, so unbuffered pushes serialise on them.
constexpr int PB_SMALL = 256;  // thread items buffered per wavefront
constexpr int PB_BIG = 64;     // wavefront-item vertices buffered per wavefront
struct WavePush {
  int ns, nb;     // buffered entries (wave-uniform)
  int jb, lo;     // owner job and list of the buffered entries (-1: none)
  int32_t *sm;    // LDS [PB_SMALL]
  int32_t *bv;    // LDS [PB_BIG] vertices
  int32_t *bn;    // LDS [PB_BIG] chunk counts
};

__device__ __forceinline__ void wp_init(WavePush &b, int32_t *lds_small, int32_t *lds_big) {
  const int w = threadIdx.x >> 6;
  b.ns = b.nb = 0;
  b.jb = -1;
  b.lo = 0;
  b.sm = lds_small + w * PB_SMALL;
  b.bv = lds_big + w * 2 * PB_BIG;
  b.bn = b.bv + PB_BIG;
}

// Called converged by the whole wavefront.
__device__ void wp_flush(WavePush &b, const JobConst *jcs, JobState *jss, const GraphDev &g) {
  __builtin_amdgcn_wave_barrier();
  const int lane = lane_id();
  if (b.jb >= 0 && (b.ns || b.nb)) {
    const JobConst &c = jcs[b.jb];
    JobState *s = &jss[b.jb];
    // chunk offsets of the buffered hub vertices, then both ranges reserved by one
    // atomic instruction (lane 0: thread items, lane 1: wavefront items)
    const int nch = lane < b.nb ? b.bn[lane] : 0;
    int x = nch;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    const int tot = __shfl(x, 63, 64);
    int r = 0;
    if ((lane == 0 && b.ns) || (lane == 1 && b.nb)) r = atomicAdd(&s->cnt[b.lo][lane], lane ? tot : b.ns);
    const int base_s = __shfl(r, 0, 64), base_b = __shfl(r, 1, 64);
    if (b.ns) {
      gi32 *dst = as_global(c.lst + (size_t)b.lo * g.n + base_s);
      for (int i = lane; i < b.ns; i += WAVE) dst[i] = b.sm[i];
    }
    if (lane < b.nb) {
      const int32_t v = b.bv[lane];
      gi32 *dst = as_global(c.blst + ((size_t)b.lo * c.bcap + base_b + x - nch) * 2);
      for (int i = 0; i < nch; i++) { dst[2 * i] = v; dst[2 * i + 1] = i; }
    }
  }
  b.ns = b.nb = 0;
  __builtin_amdgcn_wave_barrier();
}

// Buffer u (for lanes with want) into job jb's list lo. Called converged.
__device__ __forceinline__ void wp_push(WavePush &b, const JobConst *jcs, JobState *jss, const GraphDev &g,
                                        int jb, int lo, bool want, int32_t u, int nch_u) {
  const unsigned long long wm = __ballot(want);
  if (!wm) return;
  const int first = __builtin_ctzll(wm);
  const int jbF = __shfl(jb, first, 64), loF = __shfl(lo, first, 64);
  if (!__all(!want || (jb == jbF && lo == loF))) {  // mixed jobs or lists in one step: rare
    wp_flush(b, jcs, jss, g);
    // by nch_u, as the buffered path does: far-pile pushes are thread items (nch_u = 0)
    if (want) push_item(jcs[jb], &jss[jb], g, lo, u, nch_u);
    return;
  }
  if (b.jb != jbF || b.lo != loF) {
    wp_flush(b, jcs, jss, g);
    b.jb = jbF;
    b.lo = loF;
  }
  const bool small = want && nch_u == 0;
  const bool big = want && nch_u > 0;
  const unsigned long long ms = __ballot(small), mb = __ballot(big);
  const int cs = __popcll(ms), cb = __popcll(mb);
  if (b.ns + cs > PB_SMALL || b.nb + cb > PB_BIG) wp_flush(b, jcs, jss, g);
  const unsigned lo32s = (unsigned)ms, hi32s = (unsigned)(ms >> 32);
  const unsigned lo32b = (unsigned)mb, hi32b = (unsigned)(mb >> 32);
  if (small) b.sm[b.ns + __builtin_amdgcn_mbcnt_hi(hi32s, __builtin_amdgcn_mbcnt_lo(lo32s, 0))] = u;
  if (big) {
    const int r = b.nb + __builtin_amdgcn_mbcnt_hi(hi32b, __builtin_amdgcn_mbcnt_lo(lo32b, 0));
    b.bv[r] = u;
    b.bn[r] = nch_u;
  }
  b.ns += cs;
  b.nb += cb;
}

// --------------------------------------------------------------------------- select
// Non-hop: evaluate every frontier vertex's active states once per round (update UB,
// apply the UB/2 and lower-bound cuts) and overwrite its mask with the states that grow
// now. States above the near-far threshold move to the far pile. Hub vertices spread
// over many wavefront items are evaluated once, not per item.
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_select(const JobConst *jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long so0[MAXJOBS + 1], so1[MAXJOBS + 1];
  __shared__ int32_t pb_small[(BLK / WAVE) * PB_SMALL];
  __shared__ int32_t pb_big[(BLK / WAVE) * 2 * PB_BIG];
  WavePush pb;
  wp_init(pb, pb_small, pb_big);
  const int nj = load_plan(pl, WK_THR, so0);
  load_plan(pl, WK_WAVE, so1);
  const long long t0 = so0[nj], total = t0 + so1[nj];
  GrowCounters k;
  k.zero();
  int cj = -1;
  const long long base0 = (long long)blockIdx.x * BLK + (threadIdx.x & ~63);
  for (long long wbase = base0; wbase < total; wbase += (long long)gridDim.x * BLK) {
    const long long it = wbase + lane_id();
    int jb = -1, fl = 0;
    int32_t v = 0;
    bool want = false;
    if (it < total) {
      const int32_t *item;
      if (it < t0) {
        jb = find_job(so0, nj, it);
        const JobConst &c = jcs[jb];
        item = c.lst + (size_t)cur_list(c, &jss[jb]) * g.n + (it - so0[jb]);
        v = item[0];
      } else {
        const long long i1 = it - t0;
        jb = find_job(so1, nj, i1);
        const JobConst &c = jcs[jb];
        item = c.blst + ((size_t)cur_list(c, &jss[jb]) * c.bcap + (i1 - so1[jb])) * 2;
        v = item[0];
        if (item[1] != 0) jb = -1;  // one evaluation per vertex
      }
      if (jb >= 0) {
        if (jb != cj) { flush_lane(jcs, jss, cj, k); k.zero(); cj = jb; }
        const JobConst &c = jcs[jb];
        JobState *s = &jss[jb];
        unsigned long long *act = c.act + (size_t)cur_list(c, s) * g.n;
        const unsigned long long a = act[v];
        if (c.hop) {
          const unsigned long long gm = grow_select_hop(c, s, v, a, s->level, s->depth, load_ub(s), k.states);
          if (gm != a) act[v] = gm;
        } else {
          unsigned long long defer = 0;
          const unsigned long long gm =
              grow_select(c, s, v, a, s->level, load_ub(s), k.states, s->theta, defer, k.dmin);
          if (gm != a) act[v] = gm;
          fl = FAR_LIST + s->fpar;
          if (defer) want = atomicOr(&c.act[(size_t)FAR_ACT * g.n + v], defer) == 0;
        }
      }
    }
    wp_push(pb, jcs, jss, g, jb, fl, want, v, 0);
  }
  wp_flush(pb, jcs, jss, g);
  flush_counters(jcs, jss, cj, k);
}

// Near-far: the near frontier is empty and the threshold was raised; move the far
// states now under it to the next near list, the rest to the other far pile.
__global__ __launch_bounds__(BLK) void k_farsplit(const JobConst *jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int32_t pb_small[(BLK / WAVE) * PB_SMALL];
  __shared__ int32_t pb_big[(BLK / WAVE) * 2 * PB_BIG];
  WavePush pb;
  wp_init(pb, pb_small, pb_big);
  const int nj = load_plan(pl, WK_FAR, soff);
  const long long total = soff[nj];
  GrowCounters k;
  k.zero();
  int cj = -1;
  const long long base0 = (long long)blockIdx.x * BLK + (threadIdx.x & ~63);
  for (long long wbase = base0; wbase < total; wbase += (long long)gridDim.x * BLK) {
    const long long it = wbase + lane_id();
    int jb = -1, lo = 0, fl = 0, nch = 0;
    int32_t v = 0;
    bool wantN = false, wantF = false;
    if (it < total) {
      jb = find_job(soff, nj, it);
      if (jb != cj) { flush_lane(jcs, jss, cj, k); k.zero(); cj = jb; }
      const JobConst &c = jcs[jb];
      JobState *s = &jss[jb];
      unsigned long long *far = c.act + (size_t)FAR_ACT * g.n;
      v = c.lst[(size_t)(FAR_LIST + s->fpar) * g.n + (it - soff[jb])];
      const unsigned long long a = far[v];
      far[v] = 0;
      lo = s->parity ^ 1;
      fl = FAR_LIST + (s->fpar ^ 1);
      nch = chunks_of(c, v);
      unsigned long long defer = 0;
      int dummy = 0;
      const unsigned long long near = grow_select(c, s, v, a, s->level, load_ub(s), dummy,
                                                  s->theta, defer, k.dmin);
      if (near) wantN = atomicOr(&c.act[(size_t)lo * g.n + v], near) == 0;
      if (defer) wantF = atomicOr(&far[v], defer) == 0;
    }
    wp_push(pb, jcs, jss, g, jb, lo, wantN, v, nch);
    wp_push(pb, jcs, jss, g, jb, fl, wantF, v, 0);
  }
  wp_flush(pb, jcs, jss, g);
  flush_counters(jcs, jss, cj, k);
}

// --------------------------------------------------------------------------- grow
// Relax up to NE edges of one lane for the subsets of G. The target states of all NE
// edges are loaded in one batch (NE * G.n loads in flight per lane); the arena atomicMin
// is fire-and-forget; the mask atomicOr (whose result decides who pushes the target) is
// skipped when the bits are already set.
// dbase/dstride: target u's level row is dbase + u * dstride. Edges with h[j] false are
// skipped (want[j] = false, nch[j] = 0).
// ub: the job's upper bound (INF when it does not prune). A relaxation whose result
// exceeds it is not made: every state of an optimal derivation costs at most OPT <= ub,
// and a state above ub can neither lower ub (every candidate built on it costs more) nor
// become a merge half of a state <= ub, so its target is not even read. n_rel counts the
// relaxations made (target loads).
// acap: the largest cost an improved target is activated at (act_cap); a state above it
// is stored, as a merge input, but the select pass would drop it unevaluated for growth,
// so it is not pushed.
template <int NE>
__device__ __forceinline__ void relax_edges(int32_t *dbase, size_t dstride, unsigned long long *act_lo,
                                            const int32_t *nchk, int32_t nheavy, const RelaxGroup &G, const bool (&h)[NE],
                                            const int32_t (&u)[NE], const int32_t (&w)[NE], bool (&want)[NE],
                                            int (&nch)[NE], int &n_atom, int ub, int acap, int &n_rel,
                                            const unsigned long long *alb) {
  gi32 *Td[NE];
  gu64 *const act = as_global(act_lo);
  int t[NE][RG];  // 0 where no relaxation is made: no candidate (>= 1) is below it
  int lj[NE];     // the largest source cost arc j relaxes (-1: none)
#pragma unroll
  for (int j = 0; j < NE; j++) lj[j] = h[j] ? ub - w[j] : -1;
  if (alb) {  // levels >= 2 of a job with alive lists: no state of a dead target is made
    int cmin = INF;  // the group's cheapest state: arcs with lj below it relax nothing
#pragma unroll
    for (int i = 0; i < RG; i++) cmin = (i < G.n && G.c[i] < cmin) ? G.c[i] : cmin;
    const gi32 *ab32 = as_global((const int32_t *)alb);  // little-endian halves of the words
    uint32_t ab[NE];
#pragma unroll
    for (int j = 0; j < NE; j++) ab[j] = lj[j] >= cmin ? (uint32_t)ab32[u[j] >> 5] : 0u;
#pragma unroll
    for (int j = 0; j < NE; j++) lj[j] = ((ab[j] >> (u[j] & 31)) & 1) ? lj[j] : -1;
  }
#pragma unroll
  for (int j = 0; j < NE; j++) {
    Td[j] = as_global(dbase + (size_t)u[j] * dstride);
#pragma unroll
    for (int i = 0; i < RG; i++) {
      const bool r = i < G.n && G.c[i] <= lj[j];
      n_rel += (int)r;
      t[j][i] = r ? Td[j][G.p[i]] : 0;
    }
  }
  unsigned long long imp[NE];
#pragma unroll
  for (int j = 0; j < NE; j++) {
    imp[j] = 0;
#pragma unroll
    for (int i = 0; i < RG; i++) {
      if (h[j] && i < G.n) {
        int nc = G.c[i] + w[j];
        nc = nc < INF ? nc : INF;
        if (nc < t[j][i]) {  // t is 0 where lj excluded the relaxation
          __hip_atomic_fetch_min(&Td[j][G.p[i]], nc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          n_atom++;
          if (nc <= acap) imp[j] |= 1ull << G.p[i];
        }
      }
    }
  }
  // masks and chunk counts are read only for targets that improved: every extra
  // per-edge load would fetch another random line
  unsigned long long cur[NE];
#pragma unroll
  for (int j = 0; j < NE; j++) cur[j] = imp[j] ? act[u[j]] : ~0ull;
#pragma unroll
  for (int j = 0; j < NE; j++)
    want[j] = (imp[j] & ~cur[j])
                  ? __hip_atomic_fetch_or(&act[u[j]], imp[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0
                  : false;
#pragma unroll
  for (int j = 0; j < NE; j++) nch[j] = want[j] ? chunks_of(nchk, nheavy, u[j]) : 0;
}

// Arcs of the row [e0, e0 + deg) with weight <= lim: rows are sorted by weight, so they
// are a prefix, found by binary search (first and last weight checked first).
__device__ __forceinline__ int arcs_within(const int32_t *wts, int32_t e0, int deg, int lim) {
  if (deg <= 0) return 0;
  const gi32 *w = as_global(wts) + e0;
  const int wf = w[0], wl = w[deg - 1];
  if (wl <= lim) return deg;
  if (wf > lim) return 0;
  int lo = 1, hi = deg - 1;  // w[lo - 1] <= lim < w[hi]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (w[mid] <= lim) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Upper bound a job's grow kernels cut relaxations at (INF when it does not prune).
__device__ __forceinline__ int relax_ub(const JobConst &c, const JobState *s) {
  return (c.flags & JF_NO_PRUNE) ? INF : load_ub(s);
}

// The alive bitmap a job's grows of level L skip dead targets by (nullptr: none): built by
// k_alive after the tables for non-hop pruned jobs. A vertex whose full-tree lower bound
// exceeds UB lies on no optimal tree, so no state of it at level >= 2 belongs to an
// optimal derivation; its merged states are never written (they keep the INF of the
// query-start fill), and with this its grown ones are not either.
#ifndef GST_ALIVE_GROW
#define GST_ALIVE_GROW 1
#endif
__device__ __forceinline__ const unsigned long long *grow_alb(const JobConst &c, const JobState *s, int L) {
  return (GST_ALIVE_GROW && c.sparse && !c.passive && L >= 2 && s->n_alive >= 0) ? c.alb : nullptr;
}

// Largest cost at which a grown state is activated: non-hop states above floor(UB/2) are
// never grown (the centroid rule of the select pass); diameter-bounded ones are grown up
// to UB (their select pass has no UB/2 rule).
__device__ __forceinline__ int act_cap(const JobConst &c, int ub, int L) {
  return (c.hop || ub >= INF) ? ub : grow_cap(ub, L);
}

// Smallest source cost of a relax group.
__device__ __forceinline__ int group_min(const RelaxGroup &G) {
  int m = INF;
#pragma unroll
  for (int i = 0; i < RG; i++) m = (i < G.n && G.c[i] < m) ? G.c[i] : m;
  return m;
}

// Thread items (degree < split), edge-balanced. A wavefront takes 64 consecutive items and
// spreads their arcs over its lanes: a wave-level prefix sum of the item degrees gives
// each item's first arc slot, and a lane finds the item of its slot by binary search in
// LDS, so no lane idles behind a longer item (degrees below the split vary several-fold).
// Each item's job fields and its group of up to RG active source states are cached in
// LDS; items with more active states take further passes. Every LDS hand-off stays inside
// one wavefront.
struct ThreadSeg {
  int32_t off[WAVE + 1];  // first arc slot of each item; off[WAVE] = arcs of this pass
  int32_t e0[WAVE];       // first CSR arc of the item
  int32_t jb[WAVE], lo[WAVE];
  int32_t gn[WAVE];       // states in this pass's group
  int32_t gp[WAVE][RG], gc[WAVE][RG];
  int32_t ub[WAVE];       // the item's job's upper bound
  int32_t acap[WAVE];     // and its activation cap
  int32_t *dbase[WAVE];
  unsigned long long *act[WAVE];
  int32_t dstride[WAVE];  // row stride H * C(k, L) <= 16 * 35
  const int32_t *nck[WAVE];  // the item's job: chunk counts and n_heavy
  int32_t nhv[WAVE];
};

__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_grow_thread(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int32_t pb_small[(BLK / WAVE) * PB_SMALL];
  __shared__ int32_t pb_big[(BLK / WAVE) * 2 * PB_BIG];
  __shared__ ThreadSeg segs[BLK / WAVE];
  WavePush pb;
  wp_init(pb, pb_small, pb_big);
  const int nj = load_plan(pl, WK_THR, soff);
  const long long total = soff[nj];
  GrowCounters k;
  k.zero();
  int cj = -1;
  const int lane = lane_id();
  ThreadSeg &S = segs[threadIdx.x >> 6];
  const long long stride = (long long)gridDim.x * BLK;
  const int32_t *__restrict__ adj = g.adj;
  const int32_t *__restrict__ wts = g.w;
  // Fields of job ucj, wave-uniform: most waves hold items of one job only, and these
  // are read once per job instead of once per pass (each read of the job state would be
  // a dependent round trip ahead of the items' loads).
  int ucj = -1, ulo = 0, uheavy = 0;
  long long urow = 0;
  const int32_t *ulst = nullptr, *usrc = nullptr, *unck = nullptr;
  int32_t *udst = nullptr;
  const unsigned long long *uacur = nullptr;
  unsigned long long *uanxt = nullptr;
  const unsigned long long *ualb = nullptr;  // the job's alive bitmap when its grows use it
  for (long long wbase = (long long)blockIdx.x * BLK + (threadIdx.x & ~63); wbase < total; wbase += stride) {
    const long long it = wbase + lane;
    int deg = 0, iub = INF;
    int32_t e0 = 0;
    const int32_t *src = nullptr;
    unsigned long long gm = 0;
    const int jf = __builtin_amdgcn_readfirstlane(find_job(soff, nj, wbase));
    const long long wl = wbase + WAVE - 1 < total ? wbase + WAVE - 1 : total - 1;
    const bool uni = jf == __builtin_amdgcn_readfirstlane(find_job(soff, nj, wl));
    if (uni && jf != ucj) {
      ucj = jf;
      const JobConst &c = jcs[jf];
      const JobState *s = &jss[jf];
      const int li = __builtin_amdgcn_readfirstlane(cur_list(c, s));
      ulo = __builtin_amdgcn_readfirstlane(nxt_list(c, s));
      const int L = __builtin_amdgcn_readfirstlane(s->level);
      const int d = c.hop ? __builtin_amdgcn_readfirstlane(s->depth) : 0;
      urow = (long long)c.H * c.nsub[L];
      ulst = c.lst + (size_t)li * g.n;
      uacur = c.act + (size_t)li * g.n;
      uanxt = c.act + (size_t)ulo * g.n;
      usrc = lrow(c, L, 0, d);
      udst = lrow(c, L, 0, c.hop ? d + 1 : 0);
      unck = c.nchunk;
      uheavy = c.n_heavy;
      ualb = grow_alb(c, s, L);
    }
    if (it < total) {
      int jb, lo;
      int32_t v;
      if (uni) {
        jb = jf;
        lo = ulo;
        v = as_global(ulst)[it - soff[jf]];
        gm = as_global(uacur)[v];  // already reduced by k_select
        src = usrc + (size_t)v * urow;
        S.dbase[lane] = udst;
        S.dstride[lane] = urow;
        S.act[lane] = uanxt;
        S.nck[lane] = unck;
        S.nhv[lane] = uheavy;
      } else {  // a wave across two jobs: per-lane fields
        jb = find_job(soff, nj, it);
        const JobConst &c = jcs[jb];
        const JobState *s = &jss[jb];
        const int li = cur_list(c, s), L = s->level;
        const int d = c.hop ? s->depth : 0;
        lo = nxt_list(c, s);
        v = c.lst[(size_t)li * g.n + (it - soff[jb])];
        gm = c.act[(size_t)li * g.n + v];
        src = lrow(c, L, v, d);
        S.dbase[lane] = lrow(c, L, 0, c.hop ? d + 1 : 0);
        S.dstride[lane] = (long long)c.H * c.nsub[L];
        S.act[lane] = c.act + (size_t)lo * g.n;
        S.nck[lane] = c.nchunk;
        S.nhv[lane] = c.n_heavy;
      }
      if (jb != cj) { flush_lane(jcs, jss, cj, k); k.zero(); cj = jb; }
      k.verts++;
      e0 = g.beg[v];
      deg = g.beg[v + 1] - e0;
      iub = relax_ub(jcs[jb], &jss[jb]);
      if (!deg) gm = 0;
      S.e0[lane] = e0;
      S.jb[lane] = jb;
      S.lo[lane] = lo;
      S.ub[lane] = iub;
      S.acap[lane] = act_cap(jcs[jb], iub, jss[jb].level);
    }
    for (;;) {  // one pass per group of up to RG active states of every item
      RelaxGroup G;
      G.n = 0;
      if (gm) take_group(src, gm, G);
      // only the row's prefix of arcs light enough for some state of the group
      const int cnt = G.n ? arcs_within(wts, e0, deg, iub - group_min(G)) : 0;
      k.edges += cnt;
      if (!__any(cnt > 0)) break;
      S.gn[lane] = G.n;
#pragma unroll
      for (int i = 0; i < RG; i++) { S.gp[lane][i] = G.p[i]; S.gc[lane][i] = G.c[i]; }
      int x = cnt;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      S.off[lane] = x - cnt;
      const int T = __shfl(x, WAVE - 1, 64);
      if (lane == WAVE - 1) S.off[WAVE] = x;
      __builtin_amdgcn_wave_barrier();
      for (int a0 = 0; a0 < T; a0 += WAVE) {  // uniform trip count
        const int a = a0 + lane;
        bool want[1] = {false};
        int32_t u[1] = {0};
        int nch[1] = {0}, ljb = -1, llo = 0;
        if (a < T) {
          int lo_ = 0, hi = WAVE - 1;  // the last item whose first slot is <= a
          while (lo_ < hi) {
            const int mid = (lo_ + hi + 1) >> 1;
            if (S.off[mid] <= a) lo_ = mid; else hi = mid - 1;
          }
          const int j = lo_;
          const int32_t e = S.e0[j] + (a - S.off[j]);
          u[0] = adj[e];
          const int32_t w[1] = {wts[e]};
          const bool h[1] = {true};
          RelaxGroup Gj;
          Gj.n = S.gn[j];
#pragma unroll
          for (int i = 0; i < RG; i++) { Gj.p[i] = S.gp[j][i]; Gj.c[i] = S.gc[j][i]; }
          relax_edges<1>(S.dbase[j], (size_t)S.dstride[j], S.act[j], S.nck[j], S.nhv[j], Gj, h, u, w, want, nch, k.atom,
                         S.ub[j], S.acap[j], k.relax, uni ? ualb : nullptr);
          ljb = S.jb[j];
          llo = S.lo[j];
        }
        wp_push(pb, jcs, jss, g, ljb, llo, want[0], u[0], nch[0]);
        k.push += (int)want[0];
      }
      __builtin_amdgcn_wave_barrier();  // the next pass rewrites the segment
    }
  }
  wp_flush(pb, jcs, jss, g);
  flush_counters(jcs, jss, cj, k);
}

// Wavefront per (vertex, edge chunk) for vertices with degree >= split. Each lane
// relaxes GW_NE edges per strip of GW_NE * 64, with all their target loads in flight
// together: most arcs of this kernel sit in chunks of 512-1024 edges, which then take
// 2-4 dependent round-trip chains instead of 4-8.
constexpr int GW_NE = 4;
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_grow_wave(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int32_t pb_small[(BLK / WAVE) * PB_SMALL];
  __shared__ int32_t pb_big[(BLK / WAVE) * 2 * PB_BIG];
  WavePush pb;
  wp_init(pb, pb_small, pb_big);
  const int nj = load_plan(pl, WK_WAVE, soff);
  const long long total = soff[nj];
  const int lane = lane_id();
  GrowCounters k;
  k.zero();
  int cj = -1;
  const int32_t *__restrict__ adj = g.adj;
  const int32_t *__restrict__ wts = g.w;
  const long long waves_per_grid = (long long)gridDim.x * (BLK / WAVE);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar registers
  // fields of job cj (set when cj changes)
  int jlo = 0, jL = 0, jd = 0, jheavy = 0, jub = INF, jacap = INF;
  long long jrow = 0;
  int32_t *jsrc = nullptr, *jdst = nullptr;
  const int32_t *jnck = nullptr;
  unsigned long long *jact = nullptr;
  const unsigned long long *jalb = nullptr;
  // The wave's items are it0, it0 + S, it0 + 2S, ... (S = waves in the grid). Each lane
  // fetches the header of one of the next 64 (vertex, chunk, active mask, arc range, job),
  // so the header loads of 64 items share one memory latency; the wave then walks them.
  for (long long base = (long long)blockIdx.x * (BLK / WAVE) + wave; base < total; base += WAVE * waves_per_grid) {
    const long long my = base + (long long)lane * waves_per_grid;
    int hjb = 0, hv = 0, heb = 0, hee = 0, hch = 0;
    unsigned long long hgm = 0;
    if (my < total) {
      hjb = find_job(soff, nj, my);
      const JobConst &c = jcs[hjb];
      const int li = cur_list(c, &jss[hjb]);
      const gi32 *item = as_global(c.blst + ((size_t)li * c.bcap + (my - soff[hjb])) * 2);
      hv = item[0];
      hch = item[1];
      hgm = as_global(c.act)[(size_t)li * g.n + hv];  // already reduced by k_select
      heb = g.beg[hv] + hch * c.chunk;
      hee = min(g.beg[hv + 1], heb + c.chunk);
    }
    const long long left = (total - base + waves_per_grid - 1) / waves_per_grid;
    const int nitems = (int)(left < WAVE ? left : WAVE);
    for (int l = 0; l < nitems; l++) {
    const int jb = __builtin_amdgcn_readlane(hjb, l);
    if (jb != cj) {
      // the job's list, level and rows: read once per job, not per item (every read of
      // the job state is a dependent round trip ahead of the item's loads)
      flush_lane(jcs, jss, cj, k);
      k.zero();
      cj = jb;
      const JobConst &c = jcs[jb];
      const JobState *s = &jss[jb];
      jlo = __builtin_amdgcn_readfirstlane(nxt_list(c, s));
      jL = __builtin_amdgcn_readfirstlane(s->level);
      jd = c.hop ? __builtin_amdgcn_readfirstlane(s->depth) : 0;
      jrow = (long long)c.H * c.nsub[jL];
      jsrc = lrow(c, jL, 0, jd);
      jdst = lrow(c, jL, 0, c.hop ? jd + 1 : 0);
      jact = c.act + (size_t)jlo * g.n;
      jnck = c.nchunk;
      jheavy = __builtin_amdgcn_readfirstlane(c.n_heavy);
      // the job's upper bound as the wave first meets the job: it only decreases, so a
      // value read earlier in the launch still bounds every optimal derivation
      jub = __builtin_amdgcn_readfirstlane(relax_ub(c, s));
      jacap = act_cap(c, jub, jL);
      jalb = grow_alb(c, s, jL);
    }
    const int lo = jlo;
    const size_t dstride = (size_t)jrow;
    int32_t *const dbase = jdst;
    unsigned long long *const act_lo = jact;
    const int32_t v = __builtin_amdgcn_readlane(hv, l), ch = __builtin_amdgcn_readlane(hch, l);
    unsigned long long gm = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(hgm >> 32), l) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hgm, l);
    const int32_t *src = jsrc + (size_t)v * dstride;
    if (lane == 0 && ch == 0) k.verts++;
    if (!gm) continue;
    const int32_t eb = __builtin_amdgcn_readlane(heb, l);
    const int32_t ee = __builtin_amdgcn_readlane(hee, l);
    const int ub = jub;
    const int w0 = wts[eb];  // the chunk's lightest arc (eb < ee: chunks are never empty)
    while (gm) {
      RelaxGroup G;
      take_group(src, gm, G);
      uniform_group(G);
      // the row is sorted by weight: arcs heavier than lim lift every state of the group
      // above the upper bound, and so does the rest of the row after the first of them
      const int lim = ub - group_min(G);
      if (w0 > lim) continue;
      for (int32_t e = eb; e < ee; e += GW_NE * WAVE) {  // uniform trip count
        bool h[GW_NE], want[GW_NE];
        int32_t u[GW_NE], w[GW_NE];
        int nch[GW_NE];
#pragma unroll
        for (int j = 0; j < GW_NE; j++) {
          const int32_t ej = e + j * WAVE + lane;
          h[j] = ej < ee;
          u[j] = h[j] ? adj[ej] : 0;
          w[j] = h[j] ? wts[ej] : 0;
        }
        if (__builtin_amdgcn_readfirstlane(w[0]) > lim) break;  // lane 0 holds arc e
        if (lane == 0) k.edges += min(GW_NE * WAVE, ee - e);
        relax_edges<GW_NE>(dbase, dstride, act_lo, jnck, jheavy, G, h, u, w, want, nch, k.atom, ub, jacap, k.relax, jalb);
#pragma unroll
        for (int j = 0; j < GW_NE; j++) {
          wp_push(pb, jcs, jss, g, jb, lo, want[j], u[j], nch[j]);
          k.push += (int)want[j];
        }
      }
    }
    }
  }
  wp_flush(pb, jcs, jss, g);
  flush_counters(jcs, jss, cj, k);
}

// Zero the active masks of the list just consumed (after both grow kernels ran), or of
// the height-D list at the end of a hop level.
__global__ __launch_bounds__(BLK) void k_clear(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  const int nj = load_plan(pl, WK_CLEAR, soff);
  const long long total = soff[nj];
  for (long long it = (long long)blockIdx.x * BLK + threadIdx.x; it < total; it += (long long)gridDim.x * BLK) {
    const int jb = find_job(soff, nj, it);
    const JobConst &c = jcs[jb];
    const JobState *s = &jss[jb];
    const int li = s->phase == PH_HOPCLEAR ? c.D : cur_list(c, s);
    const int n0 = s->cnt[li][0];
    const long long i = it - soff[jb];
    const int32_t v = i < n0 ? c.lst[(size_t)li * g.n + i] : c.blst[((size_t)li * c.bcap + (i - n0)) * 2];
    c.act[(size_t)li * g.n + v] = 0;
  }
}

// --------------------------------------------------------------------------- tables
// After level 1: r1 = UB/2 + 1, group distance lower bounds vv[a][b] and Held-Karp tour
// tables hk[a][b][X] (cheapest walk from group a through every group of X ending at
// b in X) and w2[a][X] = min_b hk[a][b][X]. One block per job.
// Also the job's progressive lower bound on the optimum (CPUNONHOP.h:436-496 bounds a
// state; this bounds the whole query): the Euler tour of an optimal tree visits every
// query group and costs twice the tree, and consecutive groups on it are joined by
// walks of at least vv, so OPT >= max_a min_b (hk[a][b][R] + vv[b][a]) / 2; and every
// pair of groups is joined inside the tree, so OPT >= max vv[a][b]. vv is a valid lower
// bound on group-to-group distances: values up to UB/2 are exact, and larger ones are
// clamped to r1 <= their true value. Diameter-bounded jobs take vv over paths of at
// most D hops (the min over heights of the singletons), which bound the tree's paths
// between groups in the same way.
__global__ __launch_bounds__(BLK) void k_tables(const JobConst *__restrict__ jcs, JobState *jss, int njobs, GraphDev g) {
  if ((int)blockIdx.x >= njobs) return;
  const JobConst &c = jcs[blockIdx.x];
  JobState *s = &jss[blockIdx.x];
  if (s->phase != PH_TABLES) return;
  __shared__ int vv[MAXK * MAXK];
  __shared__ int hk[MAXK * MAXK * MAXW];
  __shared__ int r1s;
  const int K = c.K, W = c.W;
  if (threadIdx.x == 0) {
    const int ub = load_ub(s);
    if (c.sparse && !c.passive) s->n_alive = 0;  // k_alive fills alv after this kernel
    // level-1 distances are exact up to grow_cap(UB, 1) + 1
    long long r = c.hop ? (long long)(ub >> 1) + 1 : (long long)grow_cap(ub, 1) + 1;
    r1s = (int)(r < LB_CAP ? r : LB_CAP);
    s->r1 = r1s;
  }
  for (int i = threadIdx.x; i < MAXK * MAXK; i += BLK) vv[i] = LB_CAP;
  __syncthreads();
  const int r1 = r1s;
  for (int a = 0; a < K; a++) {
    const long long b0 = g.goffs[c.gid[a]], b1 = g.goffs[c.gid[a] + 1];
    for (long long t = b0 + threadIdx.x; t < b1; t += BLK) {
      const int32_t x = g.gverts[t];
      for (int h = 0; h < c.H; h++) {  // non-hop: H = 1
        const int32_t *sing = lrow(c, 1, x, h);
        for (int b = 0; b < K; b++) {
          int d = sing[b];
          d = d < r1 ? d : r1;
          atomicMin(&vv[a * MAXK + b], d);
        }
      }
    }
  }
  __syncthreads();
  // Held-Karp by subset size: parallel over (a, b) for each X
  for (int X = 1; X < W; X++) {
    for (int ab = threadIdx.x; ab < K * K; ab += BLK) {
      const int a = ab / K, b = ab % K;
      int val = INF;
      if ((X >> b) & 1) {
        const int Y = X ^ (1 << b);
        if (Y == 0) val = vv[a * MAXK + b];
        else {
          for (int m = 0; m < K; m++)
            if ((Y >> m) & 1) {
              const int t = hk[(a * K + m) * W + Y] + vv[m * MAXK + b];
              val = t < val ? t : val;
            }
        }
      }
      hk[(a * K + b) * W + X] = val < LB_CAP ? val : LB_CAP;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int pair = 0, tour = 0;
    for (int a = 0; a < K; a++) {
      int mn = INF;
      for (int b = 0; b < K; b++) {
        pair = vv[a * MAXK + b] > pair ? vv[a * MAXK + b] : pair;
        const int t = hk[(a * K + b) * W + c.R] + vv[b * MAXK + a];
        mn = t < mn ? t : mn;
      }
      tour = mn > tour ? mn : tour;
    }
    const int lb = (tour + 1) >> 1;
    s->lb = lb > pair ? lb : pair;
  }
  int32_t *ghk = c.tab, *gw2 = c.tab + K * K * W, *gvv = gw2 + K * W;
  for (int i = threadIdx.x; i < K * K * W; i += BLK) ghk[i] = (i % W) ? hk[i] : 0;
  for (int i = threadIdx.x; i < K * W; i += BLK) {
    const int a = i / W, X = i % W;
    int best = LB_CAP;
    for (int b = 0; b < K; b++)
      if ((X >> b) & 1) best = hk[(a * K + b) * W + X] < best ? hk[(a * K + b) * W + X] : best;
    gw2[i] = X ? best : 0;
  }
  for (int i = threadIdx.x; i < K * K; i += BLK) gvv[i] = vv[(i / K) * MAXK + (i % K)];
}

// --------------------------------------------------------------------------- merge
// Division by a uniform divisor dv >= 1: for dv >= 2, x / dv is the high word of
// x * ceil(2^32 / dv), exact whenever x * dv < 2^32 (the product overshoots x / dv by less
// than x / 2^32 < 1 / dv). Merge tiles use it for x < 2^16 and dv < 2^8.
struct UDiv {
  uint32_t d, m;
  __device__ __forceinline__ explicit UDiv(uint32_t dv) : d(dv), m(dv > 1 ? 0xffffffffu / dv + 1u : 0u) {}
  __device__ __forceinline__ uint32_t div(uint32_t x) const { return d == 1 ? x : __umulhi(x, m); }
};

// LDS rows of the merge tiles are W + 1 words apart: with a stride of W (a multiple of
// the bank count for k >= 5) the lanes that read the same subset p of different rows
// all hit one bank, up to a 32-way conflict at the top levels.
__host__ __device__ constexpr int merge_row_stride(int W) { return W + 1; }

// Stages levels 1..L-1 of the tile's vertices [v0, v0 + tv) into LDS as
// rows[(vi * H + d) * WS + p], WS = merge_row_stride(W). Each level is one contiguous
// segment of its level block whose word i is row i / C(k,l), subset rank i % C(k,l); a
// segment holds tv * H * C(k,l) <= 2048 = SU * BLK words (the tile size is clamped to
// that), so every lane issues the loads of all levels (at most
// (MAXK-1) * SU) before the first LDS store: one memory latency per tile. lsub is the
// job's subset table, copied to LDS.
constexpr int SU = 8;
// vid (LDS, non-null for sparse merges): the tile's vertex ids, instead of v0 + vi (H = 1).
__device__ __forceinline__ void stage_levels(const JobConst &c, int L, long long v0, int tv, int H,
                                             const uint8_t (*lsub)[64], int32_t *rows, const int32_t *vid) {
  const int WS = merge_row_stride(c.W);
  const int t = (int)threadIdx.x;
  int32_t x[MAXK - 1][SU];
#pragma unroll
  for (int l = 1; l < MAXK; l++) {
    if (l < L) {
      const int ns = c.nsub[l], n = tv * H * ns;
      if (vid) {  // rows of scattered vertices: word i is rank i % ns of vertex vid[i / ns]
        const UDiv dn(ns);
        const int32_t *src = c.T + c.off[l];
#pragma unroll
        for (int u = 0; u < SU; u++) {
          const int i = u * BLK + t;
          const int r = (int)dn.div((uint32_t)i);
          x[l - 1][u] = i < n ? src[(size_t)vid[r < tv ? r : 0] * ns + (i - r * ns)] : INF;
        }
      } else {
        const int32_t *src = c.T + c.off[l] + v0 * H * ns;
#pragma unroll
        for (int u = 0; u < SU; u++) {
          const int i = u * BLK + t;
          x[l - 1][u] = i < n ? src[i] : INF;
        }
      }
    }
  }
  __syncthreads();  // lsub (written before the call) is visible from here
#pragma unroll
  for (int l = 1; l < MAXK; l++) {
    if (l < L) {
      const int ns = c.nsub[l], n = tv * H * ns;
      const UDiv dn(ns);
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int i = u * BLK + t;
        if (i < n) {
          const int r = (int)dn.div((uint32_t)i);
          rows[r * WS + lsub[l][i - r * ns]] = x[l - 1][u];
        }
      }
    }
  }
}

// Copies the job's subset table (8 x 64 bytes) to LDS; visible after the next barrier
// (stage_levels holds one).
__device__ __forceinline__ void load_lsub(const JobConst &c, uint8_t (*lsub)[64]) {
  if (threadIdx.x < 128) ((uint32_t *)lsub)[threadIdx.x] = ((const uint32_t *)c.sub)[threadIdx.x];
}

// Merge counters are kept per thread across the tiles of one job and flushed when the
// block moves to another job or ends: one atomic per wavefront, not one per thread and
// tile. Called by the whole block (cj is block-uniform).
__device__ __forceinline__ void flush_merge(const JobConst *jcs, JobState *jss, int cj,
                                            unsigned long long &merged, int &minc) {
  if (cj >= 0) {
    unsigned long long x = merged;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(stat_at(jcs[cj], ST_MERGE), x);
    if (threadIdx.x == 0 && minc < INF) atomicMin(&jss[cj].minc, minc);
  }
  merged = 0;
  minc = INF;
}

// Appends the nb buffered alive vertices of job bj (-1: none) to its list. Called
// converged by the whole wavefront.
constexpr int ALIVE_BUF = 512;
__device__ __forceinline__ void alive_flush(const JobConst *jcs, JobState *jss, int bj, const int32_t *buf, int nb) {
  __builtin_amdgcn_wave_barrier();
  if (bj >= 0 && nb) {
    const int lane = lane_id();
    int base = 0;
    if (lane == 0) base = atomicAdd(&jss[bj].n_alive, nb);
    base = __shfl(base, 0, 64);
    int32_t *dst = jcs[bj].alv + base;
    for (int i = lane; i < nb; i += WAVE) dst[i] = buf[i];
  }
  __builtin_amdgcn_wave_barrier();
}

// Vertices that can lie on an optimal tree, after the tables of a non-hop job with
// sparse merges (the same round, after k_tables): an optimal tree through v costs at
// least lb_tail(v, every group), the tour and one-label bounds over v's group distances
// (a closed walk from v around the tree visits every group), so v is alive when that
// bound is within UB. The distances are exact up to UB (level 1 grows to UB - 1), which
// makes the bound strong: on C4 at k = 5 it leaves 13-35% of the vertices. Alive vertices
// are appended to alv, and every vertex's bit of alb is written (set when alive).
// The plan pads each job to a multiple of 64 vertices, so a wavefront (64 consecutive
// items from a multiple of 64) tests 64 vertices of one job, one word of its alb.
__global__ __launch_bounds__(BLK) void k_alive(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  // per wavefront: the bound's tables for X = every group (the same for all vertices of a
  // job): hk[a][b][R] at [a * MAXK + b], w2[a][R] at [MAXK * MAXK + a]; the UB and r1
  __shared__ int tabs[BLK / WAVE][MAXK * MAXK + MAXK + 2];
  // per wavefront: alive vertices of job bj not yet appended to its list. One job counter
  // takes every append of the grid, so appends are batched to one atomicAdd per flush.
  __shared__ int32_t abuf[BLK / WAVE][ALIVE_BUF];
  const int nj = load_plan(pl, WK_ALIVE, soff);
  const long long total = soff[nj];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  int *T = tabs[wv];
  int32_t *buf = abuf[wv];
  int cached = -1, bj = -1, nb = 0;  // wave-uniform
  for (long long wbase = (long long)blockIdx.x * BLK + (threadIdx.x & ~63); wbase < total;
       wbase += (long long)gridDim.x * BLK) {
    const int jf = __builtin_amdgcn_readfirstlane(find_job(soff, nj, wbase));
    const JobConst &c = jcs[jf];
    const int K = c.K;
    if (jf != cached) {  // (re)load the tables of jf
      __builtin_amdgcn_wave_barrier();  // the previous job's tables have been read
      const int W = c.W, R = c.R;
      if (lane < MAXK * MAXK) {
        const int a = lane / MAXK, b = lane % MAXK;
        T[lane] = (a < K && b < K) ? c.tab[(a * K + b) * W + R] : 0;
      } else if (lane < MAXK * MAXK + MAXK) {
        const int a = lane - MAXK * MAXK;
        T[lane] = a < K ? c.tab[K * K * W + a * W + R] : 0;
      } else if (lane == MAXK * MAXK + MAXK) {
        T[lane] = load_ub(&jss[jf]);
      } else if (lane == MAXK * MAXK + MAXK + 1) {
        T[lane] = jss[jf].r1;
      }
      __builtin_amdgcn_wave_barrier();
      cached = jf;
    }
    const int32_t v0 = (int32_t)(wbase - soff[jf]);  // a multiple of 64
    const int32_t v = v0 + lane;
    bool alive = false;
    if (v < g.n) {
      // lb_tail(v, every group), with the tables from LDS
      const int32_t *sing = c.T + c.off[1] + (size_t)v * K;  // level-1 row (H = 1)
      const int r1 = T[MAXK * MAXK + MAXK + 1];
      int dl[MAXK];
      int one = 0, mind = INF;
#pragma unroll
      for (int i = 0; i < MAXK; i++) {
        int x = i < K ? sing[i] : 0;
        x = x < r1 ? x : r1;
        x = x < LB_CAP ? x : LB_CAP;
        dl[i] = x;
        if (i < K) { one = x > one ? x : one; mind = x < mind ? x : mind; }
      }
      int lb1 = INF, lb2 = 0;
#pragma unroll
      for (int a = 0; a < MAXK; a++) {
        if (a >= K) continue;
#pragma unroll
        for (int b = 0; b < MAXK; b++) {
          if (b >= K) continue;
          const int t = (dl[a] + T[a * MAXK + b] + dl[b]) >> 1;
          lb1 = t < lb1 ? t : lb1;
        }
        const int t2 = (dl[a] + T[MAXK * MAXK + a] + mind) >> 1;
        lb2 = t2 > lb2 ? t2 : lb2;
      }
      int lb = lb1 > lb2 ? lb1 : lb2;
      lb = lb > one ? lb : one;
      alive = lb <= T[MAXK * MAXK + MAXK];
    }
    const unsigned long long m = __ballot(alive);
    if (v0 < g.n && lane == 0) c.alb[v0 >> 6] = m;
    if (bj != jf || nb + WAVE > ALIVE_BUF) {
      alive_flush(jcs, jss, bj, buf, nb);
      bj = jf;
      nb = 0;
    }
    if (alive) buf[nb + __popcll(m & ((1ull << lane) - 1))] = v;
    nb += __popcll(m);
  }
  alive_flush(jcs, jss, bj, buf, nb);
}

// Non-hop merge for level L: T(v,p) = min over splits p = p1 + p2 of T(v,p1) + T(v,p2).
// A tile of vertices has its rows staged in LDS (coalesced contiguous read), each
// thread evaluates (vertex, subset) items, new states are activated for growth.
// LDS of a merge tile: both merge kernels hold 5 blocks per CU by their registers, so a
// block may take 32 KB of the CU's 160 KB (5888 words: 45 rows of 129 words at k = 7, 178
// of 33 at k = 5) without losing a block
constexpr int MERGE_LDS_WORDS = 5888;
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(5))) void k_merge(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int32_t rows[MERGE_LDS_WORDS];
  __shared__ unsigned long long amask[512];
  __shared__ int ubs, minc_s;
  __shared__ uint32_t lsubw[128];
  __shared__ int32_t tvid[512];  // sparse merges: the tile's vertices (entries of alv)
  uint8_t (*lsub)[64] = (uint8_t (*)[64])lsubw;
  const int nj = load_plan(pl, WK_MERGE, soff);
  const long long total = soff[nj];
  int cj = -1, minc_acc = INF;
  unsigned long long merged_acc = 0;
  for (long long gt = blockIdx.x; gt < total; gt += gridDim.x) {
  const int jb = find_job(soff, nj, gt);
  if (jb != cj) { flush_merge(jcs, jss, cj, merged_acc, minc_acc); cj = jb; }
  const JobConst &c = jcs[jb];
  JobState *s = &jss[jb];
  if (c.hop) continue;
  const int W = c.W, K = c.K, L = s->level;
  const int TV = c.tv;
  const int nL = c.nsub[L];
  const int lo = s->parity ^ 1;
  const bool prune = !(c.flags & JF_NO_PRUNE);
  const int32_t *hk = c.tab;
  const int32_t *w2 = c.tab + K * K * W;
  const int r1 = s->r1;
  int merged = 0;
  {
    // Sparse merges (after k_alive): the tile is TV consecutive entries of alv, the
    // vertices whose full-tree lower bound is within UB; every other vertex lies on no
    // optimal tree and keeps the INF its level was filled with at query start.
    const bool sp = c.sparse && !c.passive && s->n_alive >= 0;
    const long long nv = sp ? s->n_alive : g.n;
    const long long v0 = (gt - soff[jb]) * TV;
    const int tv = (int)min((long long)TV, nv - v0);
    if (sp) {
      for (int i = threadIdx.x; i < tv; i += BLK) tvid[i] = c.alv[v0 + i];
      __syncthreads();
    }
    // stage the lower levels of the tile's rows into LDS, indexed by subset p
    // every subset read below has fewer than L members, so staging overwrites all of it
    for (int i = threadIdx.x; i < tv; i += BLK) amask[i] = 0;
    if (threadIdx.x == 0) { ubs = load_ub(s); minc_s = INF; }
    load_lsub(c, lsub);
    stage_levels(c, L, v0, tv, 1, lsub, rows, sp ? tvid : nullptr);
    __syncthreads();
    const int ub = ubs;
    int ubl = ub, mact = INF;
    const UDiv dnl(nL);
    const int nsplit = 1 << (L - 1);  // splits per subset of the level, plus one
    for (int it = threadIdx.x; it < tv * nL; it += BLK) {
      const int vi = (int)dnl.div((uint32_t)it), j = it - vi * nL;
      const int p = lsub[L][j];
      const int32_t *row = rows + vi * merge_row_stride(W);
      const int low = p & -p, rest = p ^ low;
      // splits p1 + p2 = p with low in p1: sm runs over the proper subsets of rest, the
      // same count for every item of the level (a uniform loop)
      int m = INF;
      int sm = rest;
      for (int t = 1; t < nsplit; t++) {
        sm = (sm - 1) & rest;
        const int p1 = low | sm;
        m = min(m, row[p1] + row[p ^ p1]);
      }
      // every state of the level is written here, INF included: levels >= 2 are not
      // filled at query start
      const long long v = sp ? (long long)tvid[vi] : v0 + vi;
      if (c.passive && L < K)
        as_global(c.cT)[(size_t)v * cohort_stride(c.cB, nL) + c.cq * nL + j] = m < INF ? m : INF;
      else
        as_global(c.T)[(size_t)c.off[L] + (size_t)v * nL + j] = m < INF ? m : INF;
      if (m >= INF) continue;
      if (m >= SAT_AT) flag_saturation(s);
      merged++;
      int sl[MAXK];
#pragma unroll
      for (int i = 0; i < MAXK; i++) sl[i] = (i < K) ? row[1 << i] : INF;
      const int q = c.R ^ p;
      const int cand = (L == K) ? m : ub_candidate(sl, p, m, L, c, (q && popc(q) < L) ? row[q] : INF);
      if (cand < ubl) ubl = cand;
      if (L < K) {
        bool grow = true;
        if (prune) grow = m <= (ub >> 1) && m + lb_tail(sl, r1, p, c, hk, w2) <= ub;
        if (grow) {
          atomicOr(&amask[vi], 1ull << j);
          mact = m < mact ? m : mact;
        }
      }
    }
    if (ubl < ub) atomicMin(&ubs, ubl);
    if (mact < INF) atomicMin(&minc_s, mact);
    __syncthreads();
    if (threadIdx.x == 0 && ubs < ub) atomicMin(&s->ub, ubs);
    if (threadIdx.x == 0) minc_acc = minc_s < minc_acc ? minc_s : minc_acc;
    if (c.passive) {  // a cohort member: activate in list 1 of the cohort's label masks
      const int S = cohort_stride(c.cB, nL), l0 = c.cq * nL, w0 = l0 >> 6, sh = l0 & 63;
      for (int vi = threadIdx.x; vi < tv; vi += BLK) {
        const unsigned long long a = amask[vi];
        if (!a) continue;
        gu64 *mw = as_global(c.cmask) + ((size_t)g.n + v0 + vi) * (S >> 6) + w0;
        __hip_atomic_fetch_or(mw, a << sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sh + nL > 64) __hip_atomic_fetch_or(mw + 1, a >> (64 - sh), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        as_global(c.cflag)[(size_t)g.n + v0 + vi] = 1;
      }
    } else {
      for (int vi = threadIdx.x; vi < tv; vi += BLK)  // k_collect builds the worklist
        if (amask[vi]) c.act[(size_t)lo * g.n + (sp ? (long long)tvid[vi] : v0 + vi)] = amask[vi];
    }
    __syncthreads();
  }
  merged_acc += merged;
  }
  flush_merge(jcs, jss, cj, merged_acc, minc_acc);
}

// Hop merge for level L. A stored hop state (v,p,d) holds the cost of a tree rooted at v
// covering p whose height is at most d (a grow step gives height exactly d; a merge may
// store a lower tree at d). Every later use only needs "at most": a merge of heights
// <= d1 and <= d2 with d1 + d2 <= D has diameter <= D, and a grow adds one. So the merge
// takes, over the staged prefix minima pm(h,q) = min_{h' <= h} T(v,q,h'),
//   T(v,p,nd) = min over splits p1 + p2 = p of
//               min(pm(nd,p1) + pm(b,p2), pm(b,p1) + pm(nd,p2)),  b = min(nd, D - nd),
// the best pair of heights (d1, d2) with max(d1, d2) <= nd and d1 + d2 <= D. The minimum
// over heights of the full states, the answer, is the same as with exact heights.
constexpr int HMERGE_LDS_WORDS = 5888;  // e.g. 7 vertices x 6 heights x 129 words (k = 7, D = 5)
constexpr int HMERGE_MASKS = 512;       // (vertex, height) masks of a tile: tv * H is clamped to it
// Registers that carry the next tile's rows while the block evaluates the current one:
// the tile size is clamped so that a level segment holds at most SUH * BLK words (1536,
// e.g. k = 4, D = 0, level 2: 256 vertices x 6 subsets), so SUH loads per lane cover it.
constexpr int SUH = 6;

// Loads levels lo..hi-1 of the rows of vertices [v0, v0 + tv) of job c into x (word i of
// level l's contiguous segment in x[l - 1][i / BLK] of thread i % BLK), without waiting.
__device__ __forceinline__ void load_hop_tile(const JobConst &c, int lo, int hi, long long v0, int tv,
                                              int32_t (&x)[MAXK - 1][SUH]) {
  const int t = (int)threadIdx.x;
  // the subset counts as one 8-byte scalar load
  const unsigned long long nsw = *(const unsigned long long *)c.nsub;
#pragma unroll
  for (int l = 1; l < MAXK; l++) {
    if (l >= lo && l < hi) {
      const int ns = (int)((nsw >> (8 * l)) & 0xff), n = tv * c.H * ns;
      const gi32 *src = as_global(c.T + c.off[l] + v0 * c.H * ns);
      // every lane loads (past the segment end: its last word, never stored): a
      // register written by an ALU op while an older load into it may be pending would
      // make the compiler wait for all loads in flight, the previous level's included
#pragma unroll
      for (int u = 0; u < SUH; u++) x[l - 1][u] = src[min(u * BLK + t, n - 1)];
    }
  }
}

// Issues the loads of merge tile gt into x if it belongs to a hop job; returns gt, or -1.
// Bound-first tiles (BF) fetch level 1 only. The job index is made uniform and the job
// table is read-only in the kernel, so the job's fields come through scalar loads and
// never wait behind the loads in flight. lvl: each job's level, in LDS.
template <bool BF>
__device__ __forceinline__ long long issue_hop_tile(const JobConst *__restrict__ jcs, const int *lvl,
                                                    const long long *soff, int nj, long long gt, long long total,
                                                    int n, int32_t (&x)[MAXK - 1][SUH]) {
  if (gt >= total) return -1;
  const int jb = __builtin_amdgcn_readfirstlane(find_job(soff, nj, gt));
  const JobConst &c = jcs[jb];
  if (!c.hop) return -1;
  const int tvx = BF ? c.tvb : c.tv;
  const long long v0 = (gt - soff[jb]) * tvx;
  load_hop_tile(c, 1, BF ? 2 : lvl[jb], v0, (int)min((long long)tvx, (long long)n - v0), x);
  return gt;
}

// BF = false: the tiles of work kind WK_MERGE (levels 2, 3 and k, and every level when
// pruning is off); BF = true: the bound-first tiles (WK_MERGEB), compiled for 4 waves per
// SIMD (its own register needs are larger, and it must not lower the occupancy of the
// other form, whose tiles are latency-bound).
template <bool BF>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(BF ? 4 : 1))) void k_merge_hop(
    const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int lvl[MAXJOBS];
  __shared__ int32_t pm[HMERGE_LDS_WORDS];
  __shared__ unsigned long long amask[HMERGE_MASKS];
  __shared__ int ubs, ub0s;
  __shared__ uint32_t lsubw[128];
  __shared__ int32_t hq[BF ? BLK : 1];  // items queued for their split loops, WAVE entries per wavefront
  __shared__ unsigned live;             // subtiles of a bound-first tile with an item that passes
  uint8_t (*lsub)[64] = (uint8_t (*)[64])lsubw;
  for (int j = threadIdx.x; j < pl->njobs; j += BLK) lvl[j] = jss[j].level;
  const int nj = load_plan(pl, BF ? WK_MERGEB : WK_MERGE, soff);  // (its barrier also publishes lvl)
  const long long total = soff[nj];
  int cj = -1, minc_acc = INF;
  unsigned long long merged_acc = 0;
  // A tile's rows travel in registers while the block evaluates the previous tile: the
  // loads of tile gt + gridDim.x are issued once tile gt is in LDS.
  int32_t x[MAXK - 1][SUH];
  long long issued = issue_hop_tile<BF>(jcs, lvl, soff, nj, blockIdx.x, total, g.n, x);
  for (long long gt = blockIdx.x; gt < total; gt += gridDim.x) {
  const int jb = __builtin_amdgcn_readfirstlane(find_job(soff, nj, gt));
  if (jb != cj) { flush_merge(jcs, jss, cj, merged_acc, minc_acc); cj = jb; }
  const JobConst &c = jcs[jb];
  JobState *s = &jss[jb];
  if (!c.hop) continue;
  const int W = c.W, H = c.H, D = c.D, K = c.K, L = lvl[jb];
  const int WS = merge_row_stride(W), RW = H * WS;
  const int nL = c.nsub[L];
  int merged = 0;
  {
    if (issued != gt) issued = issue_hop_tile<BF>(jcs, lvl, soff, nj, gt, total, g.n, x);
    const int TVX = BF ? c.tvb : c.tv;
    const long long v0 = (gt - soff[jb]) * TVX;
    const int tvx = (int)min((long long)TVX, (long long)g.n - v0);  // vertices of the tile
    for (int i = threadIdx.x; i < tvx * H; i += BLK) amask[i] = 0;
    // UB known when the tile starts: UB candidates of the tile are compared with it, and
    // the growth filter of the hop select pass uses it. Read by one thread and shared
    // through LDS, so that no vector load is pending while the next tile's rows are.
    if (threadIdx.x == 0) { ubs = INF; ub0s = load_ub(s); live = 0; }
    load_lsub(c, lsub);
    __syncthreads();
    const int ub0 = ub0s;
    const bool prune = !(c.flags & JF_NO_PRUNE);
    int ubl = INF;
    // The (sub)tile being merged: vertices [v0 + vb, v0 + vb + tv); its rows of levels
    // 1..L-1 are staged in pm[(vi * H + d) * WS + q].
    int vb = 0, tv = tvx;
    // rows of levels 1..L-1 in x -> pm; every subset read below has fewer than L members,
    // so this overwrites all of it. Height-D singletons are never grown (so the select
    // pass never sees them): the staged values are checked for saturation here.
    auto stage = [&]() {
      bool satl = false;
#pragma unroll
      for (int l = 1; l < MAXK; l++) {
        if (l < L) {
          const int ns = c.nsub[l], n = tv * H * ns;
          const UDiv dn(ns);
#pragma unroll
          for (int u = 0; u < SUH; u++) {
            const int i = u * BLK + (int)threadIdx.x;
            if (i < n) {
              const int r = (int)dn.div((uint32_t)i);
              const int32_t xv = x[l - 1][u];
              satl |= xv >= SAT_AT && xv < INF;
              pm[r * WS + lsub[l][i - r * ns]] = xv;
            }
          }
        }
      }
      if (satl) flag_saturation(s);
    };
    auto prefix = [&]() {  // prefix minima over height, in place
      for (int i = threadIdx.x; i < tv * W; i += BLK) {
        const int vi = i >> c.K, q = i & (W - 1);
        if (q == 0 || popc(q) >= L) continue;
        int m = INF;
        for (int d = 0; d < H; d++) {
          int32_t *y = &pm[vi * RW + d * WS + q];
          m = *y < m ? *y : m;
          *y = m;
        }
      }
    };
    // item order (vertex, height, subset) makes the level-block writes contiguous
    const UDiv dnl(nL), dh(H);
    const int nsplit = 1 << (L - 1);  // splits per subset of the level, plus one
    gi32 *const Tl0 = as_global(c.T) + (size_t)c.off[L] + (size_t)v0 * H * nL;
    gi32 *Tl = Tl0;
    int nitems = tv * nL * H;
    // the rest of an item once its minimum over the splits is known
    auto finish_item = [&](int it, int j, int vi, int nd, int p, int m) {
      bool keep = m < INF;
      if (keep) {
        if (m >= SAT_AT) flag_saturation(s);
        const int q = c.R ^ p;
        const int32_t *pr = pm + (vi * H + (D - nd)) * WS;  // minima over heights <= D - nd
        if (q == 0) ubl = m < ubl ? m : ubl;
        else if (popc(q) < L && (long long)m + pr[q] < ubl) ubl = m + pr[q];
        if (prune && L < K) {
          // m + lb > UB: no optimal derivation passes through this state, as a growth
          // source or as a merge input, so it is not stored (it stays INF)
          int lb = 0;
          for (int i = 0; i < K; i++)
            if (!((p >> i) & 1)) lb = pr[1 << i] > lb ? pr[1 << i] : lb;
          if ((long long)m + lb > ub0) keep = false;
        }
      }
      // every state of the level is written here, INF included: levels >= 2 are not
      // filled at query start
      Tl[it] = keep ? m : INF;
      if (!keep) return;
      merged++;
      if (L < K && nd < D) atomicOr(&amask[(vb + vi) * H + nd], 1ull << j);
    };
    auto merge_item = [&](int it) {
      const int row = (int)dnl.div((uint32_t)it);   // vi * H + nd
      const int j = it - row * nL;
      const int vi = (int)dh.div((uint32_t)row), nd = row - vi * H;
      const int p = lsub[L][j];
      const int low = p & -p, rest = p ^ low;
      const int b = min(nd, D - nd);
      const int32_t *pn = pm + vi * RW + nd * WS;  // heights <= nd
      const int32_t *pb = pm + vi * RW + b * WS;   // heights <= b
      // splits p1 + p2 = p with low in p1: sm runs over the proper subsets of rest. The
      // count is the same for every item of the level, so the loop is uniform; an INF
      // half gives a sum >= INF (2 INF < 2^31), which never wins
      int m = INF;
      int sm = rest;
      for (int t = 1; t < nsplit; t++) {
        sm = (sm - 1) & rest;
        const int p1 = low | sm, p2 = p ^ p1;
        const int s12 = pn[p1] + pb[p2];
        const int s21 = pb[p1] + pn[p2];
        m = min(m, min(s12, s21));
      }
      finish_item(it, j, vi, nd, p, m);
    };
    // The last level has one subset (every group) per (vertex, height), so its tiles hold
    // few items (k = 7, D = 5: 7 vertices x 6 heights for 256 threads) of 2^(k-1) - 1
    // splits each: G lanes of a wavefront then share an item, lane r taking the splits
    // p1 = 1 | (x << 1) with x = r, r + G, ... < 2^(k-1) - 1 (every set with group 0
    // except the full one), and a minimum across the G lanes.
    auto merge_full = [&]() {
      int G = 2;
      while (G < 16 && nitems * G * 2 <= BLK) G *= 2;
      const int gsh = __builtin_ctz(G);
      const int nx = nsplit - 1;
      const int span = ((nitems * G + WAVE - 1) / WAVE) * WAVE;  // whole wavefronts
      for (int i = threadIdx.x; i < span; i += BLK) {
        const int it = i >> gsh, r = i & (G - 1);
        int m = INF;
        if (it < nitems) {
          const int vi = (int)dh.div((uint32_t)it), nd = it - vi * H;  // nL = 1
          const int b = min(nd, D - nd);
          const int32_t *pn = pm + vi * RW + nd * WS;
          const int32_t *pb = pm + vi * RW + b * WS;
          for (int x = r; x < nx; x += G) {
            const int p1 = 1 | (x << 1), p2 = c.R ^ p1;
            m = min(m, min(pn[p1] + pb[p2], pb[p1] + pn[p2]));
          }
        }
        for (int o = G >> 1; o > 0; o >>= 1) {
          const int y = __shfl_xor(m, o, 64);
          m = y < m ? y : m;
        }
        if (it < nitems && r == 0) finish_item(it, 0, (int)dh.div((uint32_t)it), it - (int)dh.div((uint32_t)it) * H, c.R, m);
      }
    };
    if constexpr (!BF) {
      stage();
      __syncthreads();
      prefix();
      issued = issue_hop_tile<BF>(jcs, lvl, soff, nj, gt + gridDim.x, total, g.n, x);
      __syncthreads();
      if (L == 2) {  // stars: arms of at most floor(D/2) hops, prefix minima of the singletons
        for (int vi = threadIdx.x; vi < tv; vi += BLK) {
          long long star = 0;
          for (int i = 0; i < K; i++) star += pm[(vi * H + D / 2) * WS + (1 << i)];
          if (star < ubl) ubl = (int)star;
        }
      }
      if (L == K && nitems * 2 <= BLK) merge_full();
      else
        for (int it = threadIdx.x; it < nitems; it += BLK) merge_item(it);
    } else {
      // Bound-first tile (pruning on, 4 <= L < k) of up to tvb vertices. The bound of an
      // item from the singletons alone: the tree of a state (v,p,nd) holds, for each group
      // i of p, a walk of at most nd hops from v to i, and the rest of an optimal
      // derivation through it holds a walk of at most D - nd hops to each group outside p;
      // with level 1 unpruned, the prefix minima of the singletons are exactly those
      // hop-bounded distances. So
      //   max_{i in p} T1(v,i,<=nd) + max_{i not in p} T1(v,i,<=D-nd) > UB
      // implies m + lb > UB (the state is not stored) and puts every UB candidate of the
      // item above UB (its complement part is >= the second maximum): such an item is
      // written INF without its splits. sg(vi, d, i) is the prefix minimum T1(v, i, <= d).
      auto passes = [&](int it, auto sg) {
        const int row = (int)dnl.div((uint32_t)it);
        const int j = it - row * nL;
        const int vi = (int)dh.div((uint32_t)row), nd = row - vi * H;
        const int p = lsub[L][j];
        int in = 0, out = 0;
#pragma unroll 1
        for (int i = 0; i < K; i++) {
          if ((p >> i) & 1) in = max(in, sg(vi, nd, i));
          else out = max(out, sg(vi, D - nd, i));
        }
        return (long long)in + out <= ub0;
      };
      // x holds the tile's singletons only: staged compactly as sgc[(vi * H + d) * K + i]
      // and turned into prefix minima over height
      {
        const int nw = tvx * H * K;
        bool satl = false;
#pragma unroll
        for (int u = 0; u < SUH; u++) {
          const int i = u * BLK + (int)threadIdx.x;
          if (i < nw) {
            const int r = i / K, jj = i - r * K;
            const int32_t xv = x[0][u];
            satl |= xv >= SAT_AT && xv < INF;
            pm[r * K + __builtin_ctz(lsub[1][jj])] = xv;
          }
        }
        if (satl) flag_saturation(s);
      }
      __syncthreads();
      for (int i = threadIdx.x; i < tvx * K; i += BLK) {
        const int vi = i / K, gi = i - vi * K;
        int m = INF;
        for (int d = 0; d < H; d++) {
          int32_t *y = &pm[(vi * H + d) * K + gi];
          m = *y < m ? *y : m;
          *y = m;
        }
      }
      __syncthreads();
      // the items that fail are written INF; each subtile of tv vertices with an item that
      // passes (5-25% of them at levels 4 and 5, k = 7, on the power-law workloads) is
      // then staged whole and its passing items merged. While the UB is still infinite
      // every subtile is: then no answer may exist, and every staged row is checked for
      // saturation as in the other form.
      const int nsub = (tvx + c.tv - 1) / c.tv;
      {
        auto sgc = [&](int vi, int d, int i) { return pm[(vi * H + d) * K + i]; };
        unsigned lm = 0;
        for (int it = threadIdx.x; it < tvx * nL * H; it += BLK) {
          if (passes(it, sgc)) lm |= 1u << ((int)dh.div((uint32_t)dnl.div((uint32_t)it)) / c.tv);
          else Tl0[it] = INF;
        }
        if (lm) atomicOr(&live, lm);
      }
      __syncthreads();
      const unsigned lv = ub0 == INF ? (1u << nsub) - 1 : live;
      if (!lv) issued = issue_hop_tile<BF>(jcs, lvl, soff, nj, gt + gridDim.x, total, g.n, x);
      auto sgr = [&](int vi, int d, int i) { return pm[(vi * H + d) * WS + (1 << i)]; };
      for (int st = 0; st < nsub; st++) {
        if (!((lv >> st) & 1)) continue;
        vb = st * c.tv;
        tv = min(c.tv, tvx - vb);
        Tl = Tl0 + (size_t)vb * H * nL;
        nitems = tv * nL * H;
        load_hop_tile(c, 1, L, v0 + vb, tv, x);
        __syncthreads();  // the bound pass or the previous subtile is done with pm
        stage();
        __syncthreads();
        prefix();
        if (lv >> (st + 1) == 0)  // the last subtile: x is free for the next tile
          issued = issue_hop_tile<BF>(jcs, lvl, soff, nj, gt + gridDim.x, total, g.n, x);
        __syncthreads();
        // passing items queued per wavefront in LDS and evaluated 64 at a time, one per
        // lane, so that their uniform split loops run on full wavefronts
        int32_t *wq = hq + (threadIdx.x & ~(WAVE - 1));
        const int lane = lane_id();
        int pend = 0;  // queued items of this wavefront (uniform)
        for (int it0 = 0; it0 < nitems; it0 += BLK) {  // uniform trip count
          const int it = it0 + (int)threadIdx.x;
          const bool pass = it < nitems && passes(it, sgr);
          const unsigned long long pm_ = __ballot(pass);
          const int np = __popcll(pm_);
          if (pend + np > WAVE) {
            __builtin_amdgcn_wave_barrier();
            if (lane < pend) merge_item(wq[lane]);
            __builtin_amdgcn_wave_barrier();
            pend = 0;
          }
          if (pass)
            wq[pend + __builtin_amdgcn_mbcnt_hi((unsigned)(pm_ >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)pm_, 0))] = it;
          pend += np;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < pend) merge_item(wq[lane]);
        __builtin_amdgcn_wave_barrier();
      }
      tv = tvx;
    }
    if (ubl < ub0) atomicMin(&ubs, ubl);
    __syncthreads();
    if (threadIdx.x == 0 && ubs < ub0) atomicMin(&s->ub, ubs);
    for (int i = threadIdx.x; i < tvx * H; i += BLK) {  // k_collect builds the worklists
      const unsigned long long mk = amask[i];
      if (mk) {
        const int vi = (int)dh.div((uint32_t)i), nd = i - vi * H;
        as_global(c.act)[(size_t)nd * g.n + v0 + vi] = mk;
      }
    }
    __syncthreads();
  }
  merged_acc += merged;
  }
  flush_merge(jcs, jss, cj, merged_acc, minc_acc);
}

// Worklists of the vertices a merge activated. The merge kernels write the activation
// masks only; this pass scans them, 64 vertices of one list per wavefront item, and
// pushes every vertex with a nonzero mask through the wavefront push buffer, so the
// list-counter round trips are off the merge tiles' critical path. Masks outside the
// lists are zero when a merge starts (consumed masks are cleared every round), so a
// nonzero mask is exactly a merge activation.
__global__ __launch_bounds__(BLK) void k_collect(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ int32_t pb_small[(BLK / WAVE) * PB_SMALL];
  __shared__ int32_t pb_big[(BLK / WAVE) * 2 * PB_BIG];
  WavePush pb;
  wp_init(pb, pb_small, pb_big);
  const int nj = load_plan(pl, WK_COLLECT, soff);
  const long long total = soff[nj];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long nblk = ((long long)g.n + 63) / 64;
  const long long waves_per_grid = (long long)gridDim.x * (BLK / WAVE);
  for (long long it = (long long)blockIdx.x * (BLK / WAVE) + wave; it < total; it += waves_per_grid) {
    const int jb = find_job(soff, nj, it);
    const JobConst &c = jcs[jb];
    const JobState *s = &jss[jb];
    const long long r = it - soff[jb];
    const int li = c.hop ? (int)(r / nblk) : (s->parity ^ 1);
    const long long v = (c.hop ? r % nblk : r) * 64 + lane_id();
    const bool want = v < g.n && c.act[(size_t)li * g.n + v] != 0;
    const int nch = want ? chunks_of(c, (int32_t)v) : 0;
    wp_push(pb, jcs, jss, g, jb, li, want, (int32_t)v, nch);
  }
  wp_flush(pb, jcs, jss, g);
}

// --------------------------------------------------------------------------- final
// ans = min over v, d of (T(v, full, d) << 32 | v*H + d); a work item is a chunk of
// FINAL_CHUNK rows (v, d) of one job.
__global__ __launch_bounds__(BLK) void k_final(const JobConst *__restrict__ jcs, JobState *jss, const Plan *pl, GraphDev g) {
  __shared__ long long soff[MAXJOBS + 1];
  __shared__ unsigned long long bm;
  const int nj = load_plan(pl, WK_FINAL, soff);
  const long long total = soff[nj];
  for (long long gt = blockIdx.x; gt < total; gt += gridDim.x) {
    const int jb = find_job(soff, nj, gt);
    const JobConst &c = jcs[jb];
    if (threadIdx.x == 0) bm = ~0ull;
    __syncthreads();
    unsigned long long best = ~0ull;
    const int W = c.W, R = c.R;
    const long long rows = (long long)g.n * c.H;
    const long long r0 = (gt - soff[jb]) * FINAL_CHUNK;
    const long long r1 = min(rows, r0 + FINAL_CHUNK);
    for (long long i = r0 + threadIdx.x; i < r1; i += BLK) {
      const int32_t cst = c.T[(size_t)c.off[c.K] + i];  // level k holds one state per row
      if (cst < INF) {
        const unsigned long long key = ((unsigned long long)cst << 32) | (unsigned long long)i;
        best = key < best ? key : best;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(best, off, 64);
      best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(&bm, best);
    __syncthreads();
    if (threadIdx.x == 0 && bm != ~0ull) atomicMin(&jss[jb].ans, bm);
    __syncthreads();
  }
}

// --------------------------------------------------------------------------- rebuild
// Tree reconstruction from the cost arena alone. For a state of cost c a predecessor
// with cost(pred) (+ w) <= c always exists (costs only decrease after being set), so
// following any such predecessor terminates (a grow lowers the cost by >= 1, a merge
// shrinks the subset) and yields a structure of weight <= c; at the optimum it is an
// optimal tree. Choices are canonical: smallest split, then smallest edge index.
__device__ __forceinline__ bool in_group(const GraphDev &g, int grp, int32_t v) {
  long long lo = g.goffs[grp], hi = g.goffs[grp + 1];
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (g.gverts[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo < g.goffs[grp + 1] && g.gverts[lo] == v;
}

constexpr int RB_BLK = 1024;  // rebuild block: one per job, wide for hub adjacency scans
constexpr int RB_U = 8;       // arcs per thread per scan chunk
__global__ __launch_bounds__(RB_BLK) void k_rebuild(const JobConst *jcs, JobState *jss, int njobs, GraphDev g) {
  if ((int)blockIdx.x >= njobs) return;
  const JobConst &c = jcs[blockIdx.x];
  JobState *s = &jss[blockIdx.x];
  if (s->phase != PH_REBUILD) return;
  __shared__ int top, nout, err, cv, cp, cd, ccost;
  __shared__ long long found;
  const int H = c.H, W = c.W, D = c.D;
  const int RW = H * W;
  if (threadIdx.x == 0) {
    const unsigned long long a = s->ans;
    const long long idx = (long long)(a & 0xffffffffull);
    c.stk[0] = (int32_t)(idx / H);
    c.stk[1] = c.R;
    c.stk[2] = (int32_t)(idx % H);
    top = 1; nout = 0; err = 0;
  }
  __syncthreads();
  for (int step = 0; step < (1 << 20); step++) {
    if (top == 0 || err) break;
    __syncthreads();
    if (threadIdx.x == 0) {
      top--;
      cv = c.stk[3 * top]; cp = c.stk[3 * top + 1]; cd = c.stk[3 * top + 2];
      ccost = c.T[sidx(c, cv, cp, cd)];
      found = 0x7fffffffffffffffll;
    }
    __syncthreads();
    const int v = cv, p = cp, d = cd, cost = ccost;
    // seed: singleton at cost 0 (height 0) held by a member of that group
    if (popc(p) == 1 && cost == 0 && d == 0) {
      if (threadIdx.x == 0 && !in_group(g, c.gid[__builtin_ctz(p)], v)) err = 3;
      __syncthreads();
      continue;
    }
    // merge search: split index = (submask sm of rest) * H*H + d1*H + d2
    const int low = p & -p, rest = p ^ low;
    const int nsm = 1 << popc(rest);
    for (int it = threadIdx.x; it < nsm * H * H; it += RB_BLK) {
      const int k = it / (H * H), dd = it - k * H * H, d1 = dd / H, d2 = dd - d1 * H;
      // k-th submask of rest (pdep of k into rest)
      int sm = 0, kk = k;
      for (int b = 0; b < MAXK; b++)
        if ((rest >> b) & 1) { if (kk & 1) sm |= 1 << b; kk >>= 1; }
      const int p1 = low | sm;
      if (p1 == p) continue;
      if (c.hop) {
        if ((d1 > d2 ? d1 : d2) > d || d1 + d2 > D) continue;  // merged hop states: height <= d
      } else if (d1 | d2) continue;
      const long long t = (long long)c.T[sidx(c, v, p1, d1)] + c.T[sidx(c, v, p ^ p1, d2)];
      if (t <= cost) atomicMin((unsigned long long *)&found, (unsigned long long)it);
    }
    __syncthreads();
    if (found != 0x7fffffffffffffffll) {
      if (threadIdx.x == 0) {
        const int it = (int)found;
        const int k = it / (H * H), dd = it - k * H * H, d1 = dd / H, d2 = dd - d1 * H;
        int sm = 0, kk = k;
        for (int b = 0; b < MAXK; b++)
          if ((rest >> b) & 1) { if (kk & 1) sm |= 1 << b; kk >>= 1; }
        const int p1 = low | sm;
        if (top + 2 > c.scap) err = 4;
        else {
          c.stk[3 * top] = v; c.stk[3 * top + 1] = p1; c.stk[3 * top + 2] = d1; top++;
          c.stk[3 * top] = v; c.stk[3 * top + 1] = p ^ p1; c.stk[3 * top + 2] = d2; top++;
        }
      }
      __syncthreads();
      continue;
    }
    // grow search: neighbour u with T(u, p, d-1) + w(u,v) <= cost
    if (c.hop && d == 0) {
      if (threadIdx.x == 0) err = 5;
      __syncthreads();
      continue;
    }
    const int pd = c.hop ? d - 1 : 0;
    const int32_t e0 = g.beg[v], e1 = g.beg[v + 1];
    // ascending chunks of RB_U*RB_BLK arcs; the first chunk holding a predecessor has
    // the smallest index, so the scan stops there (hubs have millions of arcs). All
    // loads of a chunk are issued before the first comparison.
    for (int32_t cb = e0; cb < e1; cb += RB_U * RB_BLK) {
      int32_t uu[RB_U], ww[RB_U], tt[RB_U];
#pragma unroll
      for (int r = 0; r < RB_U; r++) {
        const int32_t e = cb + r * RB_BLK + (int32_t)threadIdx.x;
        uu[r] = e < e1 ? g.adj[e] : -1;
        ww[r] = e < e1 ? g.w[e] : 0;
      }
#pragma unroll
      for (int r = 0; r < RB_U; r++) tt[r] = uu[r] >= 0 ? c.T[sidx(c, uu[r], p, pd)] : INF;
#pragma unroll
      for (int r = 0; r < RB_U; r++) {
        const int32_t e = cb + r * RB_BLK + (int32_t)threadIdx.x;
        if (uu[r] >= 0 && (long long)tt[r] + ww[r] <= cost)
          atomicMin((unsigned long long *)&found, (unsigned long long)(e - e0));
      }
      __syncthreads();
      if (found != 0x7fffffffffffffffll) break;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (found == 0x7fffffffffffffffll) err = 6;
      else {
        const int32_t e = e0 + (int32_t)found;
        const int32_t u = g.adj[e];
        if (nout >= c.ecap || top + 1 > c.scap) err = 7;
        else {
          c.out[3 * nout] = u < v ? u : v;
          c.out[3 * nout + 1] = u < v ? v : u;
          c.out[3 * nout + 2] = g.w[e];
          nout++;
          c.stk[3 * top] = u; c.stk[3 * top + 1] = p; c.stk[3 * top + 2] = pd; top++;
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s->n_out = nout;
    if (err) s->err = err;
    else if (top != 0) s->err = 8;
  }
}

// --------------------------------------------------------------------------- stats
__global__ void k_count_finite(const int32_t *T, long long n, unsigned long long *out) {
  unsigned long long cnt = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    cnt += T[i] < INF;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, cnt);
}

}  // namespace gst
