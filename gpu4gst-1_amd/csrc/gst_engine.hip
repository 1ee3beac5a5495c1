// gst_engine.hip — host side of libgst.so: the C ABI declared in include/gst.h.
//
// The graph (CSR + group table) is uploaded once by gst_open and stays resident. A job
// is one query with its own DP arena and worklists ("slot"). gst_solve runs one job with
// host-driven rounds: after each round the host reads the job's state (one sync per
// round, like code/TrimCDP/src/DPQ.cu:512-544) and launches exactly-sized grids.
// gst_solve_batch and the serving API run many jobs with device-side phase control and
// the round loop captured in a HIP graph (gst_batch.inc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstddef>
#include <deque>
#include <map>
#include <list>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gst.h"
#include "gst_kernels.hpp"
#include "gst_cohort_kernels.hpp"

using namespace gst;
using clk = std::chrono::steady_clock;

namespace {

// Last error of the calling thread (a fixed buffer: setting it never allocates).
thread_local char g_err[512];
void set_err(const char *s) { std::snprintf(g_err, sizeof g_err, "%s", s); }
void set_err(const std::string &s) { set_err(s.c_str()); }

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::snprintf(g_err, sizeof g_err, "%s: %s", #x, hipGetErrorString(e_));      \
      return GST_DEVICE_ERROR;                                                       \
    }                                                                                \
  } while (0)

double secs(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// Device buffers of one job slot.
struct Slot {
  int K = 0, D = -1, H = 1, W = 1;
  size_t arena_words = 0;
  int32_t *T = nullptr;
  unsigned long long *act = nullptr;
  int32_t *lst = nullptr, *blst = nullptr, *tab = nullptr, *stk = nullptr, *out = nullptr;
  unsigned long long *stx = nullptr;  // striped work counters
  unsigned long long *alb = nullptr;  // non-hop: the alive bitmap of k_alive, a bit per vertex
  size_t bytes = 0;
  int nact = 0, nlst = 0, nblst = 0;  // mask arrays, thread lists, wavefront lists
  int ecap = 0;                       // tree edges the out buffer holds
};

// Non-hop: near masks 0/1 + far-pile mask 2; near lists 0/1 + far piles 2/3 (thread
// items only); wavefront lists 0/1. Hop: one of each per height.
inline void slot_shape(int D, int &nact, int &nlst, int &nblst) {
  if (D >= 0) { nact = nlst = nblst = D + 1; }
  else { nact = 3; nlst = 4; nblst = 2; }
}

// Rebuild stack entries: a derivation has k leaves (singleton seeds) and k - 1 merges, so
// the depth-first stack never holds more than k + 1 entries.
constexpr int SCAP = 1 << 12;
constexpr int ECAP = GST_MAX_TREE_EDGES;  // tree edges returned (each grow step adds one)

}  // namespace

struct gst_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  int32_t n = 0;
  int64_t m = 0;
  GraphDev g{};
  int32_t *d_beg = nullptr, *d_adj = nullptr, *d_w = nullptr, *d_gverts = nullptr;
  // grow tunings: chunk counts per vertex for each (degree split, chunk) in use
  struct TuneSet {
    int32_t split = 0, chunk = 0, n_heavy = 0;
    int32_t *d_nchunk = nullptr;  // per vertex: wavefront chunks (0 = thread item)
    long long items = 0;          // wavefront items of all heavy vertices
  };
  std::vector<TuneSet> tunes;
  int tune_k[MAXK + 1] = {0};     // per k: index into tunes
  int tile_k[MAXK + 1] = {0};     // per k: merge tile cap in vertices (0 = as many as LDS holds)
  long long *d_goffs = nullptr;
  std::vector<int64_t> h_goffs;
  std::vector<int32_t> h_gverts;    // group members, caller's ids
  std::vector<int32_t> comp;        // connected component of each vertex (caller's ids)
  std::vector<int32_t> h_deg;       // degrees in device order
  std::vector<int32_t> old_of_new;  // device vertex -> caller's id (empty: same numbering)
  std::vector<int32_t> hub_gd;      // [n_hubs][n_groups] hub-to-group distances (gst_hubs.inc)
  int n_hubs = 0;
  int32_t n_groups = 0;
  int64_t graph_bytes = 0;
  long long bcap = 0;  // wavefront items one worklist can hold
  gst_opts opts{};
  int32_t delta = 0;
  // single-query path: one slot per (K, D) shape, least recently used evicted beyond
  // plain_cap bytes (a mixed stream of shapes would otherwise reallocate its arena,
  // up to V * 2^k * (D+1) * 4 bytes, on almost every query)
  std::mutex solve_mu;
  std::list<std::pair<std::pair<int, int>, Slot>> plain;  // front = most recent
  int64_t plain_cap = 0;
  JobConst *d_jc = nullptr;  // [max_batch]
  JobState *d_js = nullptr;  // [max_batch]
  JobState *h_js = nullptr;  // pinned mirror [max_batch]
  JobConst *h_jc = nullptr;  // pinned staging of the single-query path's job constants
  unsigned long long *h_stx = nullptr;    // pinned [NSTRIPE * NSTX] counter stripes
  unsigned long long *h_count = nullptr;  // pinned finite-state count
  void *scratch = nullptr;                // this device's static scratch (gst_solve_noalloc)
  Plan *d_plan = nullptr;    // work layout of the single-query path
  unsigned long long *d_count = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // bracket the work kernels of one round
  hipEvent_t cev0 = nullptr, cev1 = nullptr;  // bracket one cohort grow launch (batch stream)
  std::mutex cstat_mu;
  gst_cohort_stats cstat{};                 // gst_get_cohort_stats
  unsigned merge_grid = 1024;               // resident blocks of k_merge on this device
  unsigned merge_hop_grid = 1024;           // resident blocks of k_merge_hop<false> on this device
  unsigned merge_hopb_grid = 1024;          // resident blocks of k_merge_hop<true>
  int max_batch = 64;
  // batched / served (gst_batch.inc)
  struct Batch;
  Batch *batch = nullptr;
};

namespace {

int64_t arena_bytes_for(int64_t n, int K, int D) {
  const int64_t H = D >= 0 ? D + 1 : 1;
  return n * ((int64_t)1 << K) * H * 4;
}

// Lays out the buffers of a slot for (K, D) from base, in 256-byte aligned pieces, with
// room for ecap tree edges; returns the bytes used. With base == nullptr it only counts.
size_t slot_carve(const gst_ctx *c, Slot &s, char *base, int K, int D, int ecap) {
  s.K = K; s.D = D;
  s.H = D >= 0 ? D + 1 : 1;
  s.W = 1 << K;
  s.ecap = ecap;
  slot_shape(D, s.nact, s.nlst, s.nblst);
  s.arena_words = (size_t)c->n * s.W * s.H;
  size_t off = 0;
  auto take = [&](size_t bytes) { char *p = base ? base + off : nullptr; off += (bytes + 255) & ~(size_t)255; return p; };
  s.T = (int32_t *)take(s.arena_words * 4);
  s.act = (unsigned long long *)take((size_t)s.nact * c->n * 8);
  s.lst = (int32_t *)take((size_t)s.nlst * c->n * 4);
  s.blst = (int32_t *)take((size_t)s.nblst * c->bcap * 8);
  s.tab = (int32_t *)take((size_t)(MAXK * MAXK * MAXW + MAXK * MAXW + MAXK * MAXK) * 4);
  s.stk = (int32_t *)take((size_t)SCAP * 12);
  s.out = (int32_t *)take((size_t)ecap * 12);
  s.stx = (unsigned long long *)take((size_t)NSTRIPE * NSTX * 8);
  s.alb = D < 0 ? (unsigned long long *)take((size_t)(c->n + 63) / 64 * 8) : nullptr;
  s.bytes = off;
  return off;
}

// bytes of a slot for (K, D)
size_t slot_bytes(const gst_ctx *c, int K, int D, int ecap = ECAP) {
  Slot s;
  return slot_carve(c, s, nullptr, K, D, ecap);
}

void slot_free(Slot &s) {
  if (s.T) (void)hipFree(s.T);
  s = Slot();
}

int slot_alloc(gst_ctx *c, Slot &s, int K, int D) {
  slot_free(s);
  char *base = nullptr;
  const size_t bytes = slot_bytes(c, K, D);
  hipError_t e = hipMalloc((void **)&base, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_err(std::string("hipMalloc(slot) failed: ") + hipGetErrorString(e));
    s = Slot();
    return GST_OVER_BUDGET;
  }
  slot_carve(c, s, base, K, D, ECAP);
  return GST_OK;
}

// Bring the single-query slot of shape (K, D) to the front of the cache, allocating it
// (and evicting the least recently used slots beyond plain_cap) when absent.
int plain_slot(gst_ctx *c, int K, int D) {
  for (auto it = c->plain.begin(); it != c->plain.end(); ++it)
    if (it->first == std::make_pair(K, D)) {
      c->plain.splice(c->plain.begin(), c->plain, it);
      return GST_OK;
    }
  const int64_t need = (int64_t)slot_bytes(c, K, D);
  int64_t held = 0;
  for (auto &e : c->plain) held += (int64_t)e.second.bytes;
  while (!c->plain.empty() && held + need > c->plain_cap) {
    held -= (int64_t)c->plain.back().second.bytes;
    slot_free(c->plain.back().second);
    c->plain.pop_back();
  }
  Slot s;
  int rc = slot_alloc(c, s, K, D);
  if (rc != GST_OK && !c->plain.empty()) {  // device memory short: drop the cache, retry once
    for (auto &e : c->plain) slot_free(e.second);
    c->plain.clear();
    rc = slot_alloc(c, s, K, D);
  }
  if (rc != GST_OK) return rc;
  c->plain.emplace_front(std::make_pair(K, D), s);
  return GST_OK;
}

// Validated, de-duplicated query (fixed capacity: no heap use).
// Non-hop pruned merges of levels >= 2 over the vertices k_alive lists (0: every vertex).
#ifndef GST_SPARSE_MERGE
#define GST_SPARSE_MERGE 1
#endif

struct Query {
  int32_t gids[MAXK] = {0};
  int K = 0, D = -1, flags = 0;
};

int check_query(const gst_ctx *c, const int32_t *gids, int32_t k, int32_t D, int32_t flags, Query &q) {
  if (!gids || k < 1) { set_err("query needs k >= 1 group ids"); return GST_BAD_INPUT; }
  q.K = 0;
  for (int i = 0; i < k; i++) {
    if (gids[i] < 0 || gids[i] >= c->n_groups) { set_err("group id out of range"); return GST_BAD_INPUT; }
    if (std::find(q.gids, q.gids + q.K, gids[i]) != q.gids + q.K) continue;
    if (q.K == MAXK) { set_err("k > GST_MAX_K"); return GST_BAD_INPUT; }
    q.gids[q.K++] = gids[i];
  }
  if (D > GST_MAX_D) { set_err("D > GST_MAX_D"); return GST_BAD_INPUT; }
  q.D = D < 0 ? -1 : D;
  q.flags = flags | (c->opts.flags & ~(GST_O_KEEP_ORDER | GST_O_NO_HUB_INDEX));
  return GST_OK;
}

void fill_jobconst(const gst_ctx *c, const Slot &s, const Query &q, JobConst &jc) {
  std::memset(&jc, 0, sizeof(jc));
  jc.T = s.T; jc.stx = s.stx; jc.act = s.act; jc.lst = s.lst; jc.blst = s.blst; jc.tab = s.tab;
  jc.stk = s.stk; jc.out = s.out;
  jc.bcap = c->bcap; jc.scap = SCAP; jc.ecap = s.ecap;
  jc.K = q.K; jc.D = q.D; jc.H = q.D >= 0 ? q.D + 1 : 1; jc.W = 1 << q.K; jc.R = jc.W - 1;
  jc.hop = q.D >= 0; jc.flags = q.flags; jc.nl = s.nlst;
  // non-hop pruned merges of levels >= 2 cover only the vertices k_alive lists, kept in
  // the arena's unused tail (levels 1..k take n * (W - 1) of its n * W words)
  jc.sparse = GST_SPARSE_MERGE && q.D < 0 && !(q.flags & GST_F_NO_PRUNE);
  jc.alv = s.T + (size_t)c->n * (jc.W - 1);
  jc.alb = s.alb;
  jc.delta = q.D >= 0 ? 0 : c->delta;
  {
    const gst_ctx::TuneSet &ts = c->tunes[c->tune_k[q.K]];
    jc.nchunk = ts.d_nchunk;
    jc.split = ts.split;
    jc.chunk = ts.chunk;
    jc.n_heavy = ts.n_heavy;
  }
  for (int i = 0; i < q.K; i++) jc.gid[i] = q.gids[i];
  long long members = 0;
  for (int i = 0; i < q.K; i++) members += c->h_goffs[q.gids[i] + 1] - c->h_goffs[q.gids[i]];
  jc.seed_total = (int32_t)members;
  const int ws = merge_row_stride(jc.W);
  // hop tiles also keep each level segment (tv * H * C(k,l) words) within the SUH loads
  // per lane that carry the next tile in registers
  int cmax = 1;
  for (int l = 1, cb = q.K; l <= q.K; cb = cb * (q.K - l) / (l + 1), l++) cmax = std::max(cmax, cb);
  jc.tv = jc.hop ? std::max(1, std::min({256, HMERGE_LDS_WORDS / (jc.H * ws), HMERGE_MASKS / jc.H,
                                         SUH * BLK / (jc.H * cmax)}))
                 : std::min({512, MERGE_LDS_WORDS / ws, SU * BLK / cmax});
  if (c->tile_k[q.K] > 0) jc.tv = std::min(jc.tv, c->tile_k[q.K]);
  // bound-first hop tiles (levels 4..k-1, pruned): whole subtiles of tv vertices whose
  // singletons (H * k words per vertex) fit the registers that carry a tile's level 1, the
  // merge LDS and the activation masks; at most 16 subtiles
  jc.tvb = jc.tv;
  if (jc.hop) {
    const int cap = std::min({SUH * BLK, HMERGE_LDS_WORDS} ) / (jc.H * q.K);
    jc.tvb = jc.tv * std::max(1, std::min({16, cap / jc.tv, HMERGE_MASKS / jc.H / jc.tv}));
  }
  int cnt[8] = {0};
  for (int p = 0; p < jc.W; p++) {
    const int L = __builtin_popcount(p);
    jc.rank[p] = (uint8_t)cnt[L];
    jc.sub[L][cnt[L]++] = (uint8_t)p;
  }
  for (int L = 0; L <= q.K; L++) jc.nsub[L] = (uint8_t)cnt[L];
  // level-major arena: level L (1..k) block of n * H * C(k,L) states
  long long off = 0;
  for (int L = 0; L <= 8; L++) {
    jc.off[L] = off;
    if (L >= 1 && L <= q.K) off += (long long)c->n * jc.H * cnt[L];
  }
}

void init_jobstate(JobState &js) {
  std::memset(&js, 0, sizeof(js));
  js.phase = PH_INIT;
  js.ub = INF;
  js.r1 = LB_CAP;
  js.ans = ~0ull;
  js.theta = INF;
  js.far_min = INF;
  js.minc = INF;
  js.n_alive = -1;
}

// Grid of the select and grow launches: one wave of resident blocks. Fewer waves means
// fewer end-of-launch flushes onto the shared worklist counters.
#ifndef GROW_GRID_CAP
#define GROW_GRID_CAP 1024
#endif

inline unsigned grid_for(long long items, long long per_block, unsigned cap) {
  long long b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// Launch the init kernels of a job on stream st: level 1 of the arena = INF (the merge of
// each level L >= 2 writes every state of its block, INF included), masks = 0.
void launch_job_init(hipStream_t st, const Slot &s, const gst_ctx *c, const JobConst &jc) {
  // level 1 only (every merge writes all the states of its level), or levels 1..k when
  // the merges are sparse (a vertex they skip keeps INF)
  const long long words = jc.sparse ? (long long)c->n * (jc.W - 1)
                                    : std::min<long long>((long long)s.arena_words, jc.off[2]);
  k_fill_i32<<<grid_for(words / 4, BLK, 8192), BLK, 0, st>>>(s.T, words, INF);
  k_fill_u64<<<grid_for((long long)s.nact * c->n, BLK, 4096), BLK, 0, st>>>(s.act, (long long)s.nact * c->n, 0ull);
  k_fill_u64<<<grid_for(NSTRIPE * NSTX, BLK, 64), BLK, 0, st>>>(s.stx, (long long)NSTRIPE * NSTX, 0ull);
}

// Sum the counter stripes of a finished job into its state record.
void fold_stats(const unsigned long long *stripes, JobState &js) {
  unsigned long long t[NSTX] = {0};
  for (int r = 0; r < NSTRIPE; r++)
    for (int i = 0; i < NSTX; i++) t[i] += stripes[r * NSTX + i];
  for (int i = 0; i < 8; i++) js.st[i] = t[i];
  for (int L = 0; L < 8; L++) {
    js.stl[L][0] = t[SX_LEVEL + 2 * L];
    js.stl[L][1] = t[SX_LEVEL + 2 * L + 1];
  }
}

// Some connected component holds a member of every query group (the non-hop
// feasibility condition; a diameter bound can only make a feasible query infeasible).
bool comp_feasible(const gst_ctx *c, const Query &q) {
  auto has_member_in = [&](int grp, int32_t comp) {
    for (int64_t t = c->h_goffs[grp]; t < c->h_goffs[grp + 1]; t++)
      if (c->comp[c->h_gverts[t]] == comp) return true;
    return false;
  };
  for (int64_t t = c->h_goffs[q.gids[0]]; t < c->h_goffs[q.gids[0] + 1]; t++) {
    const int32_t cc = c->comp[c->h_gverts[t]];
    bool all = true;
    for (int i = 1; i < q.K && all; i++) all = has_member_in(q.gids[i], cc);
    if (all) return true;
  }
  return false;
}

// Fills out from a finished job. ebuf holds n_raw tree edges (u, v, w) in device ids; they
// are rewritten in place as the distinct edges in the caller's ids, sorted, and out->edges
// takes ebuf when own_edges (freed by gst_result_free), else stays null (the caller's
// buffer of gst_solve_noalloc).
void finish_result(const gst_ctx *c, const Query &q, const JobState &js, int32_t *ebuf, int n_raw, bool own_edges,
                   gst_result *out) {
  std::memset(out, 0, sizeof(*out));
  out->k = q.K;
  out->D = q.D;
  out->rounds = js.rounds;
  out->grow_rounds = js.dp_rounds;
  out->states_processed = (int64_t)js.st[ST_STATES];
  out->grow_vertices = (int64_t)js.st[ST_VERTS];
  out->grow_edges = (int64_t)js.st[ST_EDGES];
  out->grow_relax = (int64_t)js.st[ST_RELAX];
  out->grow_pushes = (int64_t)js.st[ST_PUSH];
  out->arena_bytes = arena_bytes_for(c->n, q.K, q.D);
  for (int L = 0; L < 8; L++) {
    out->level_edges[L] = (int64_t)js.stl[L][0];
    out->level_relax[L] = (int64_t)js.stl[L][1];
  }
  out->root = -1;
  out->root_height = -1;
  out->level = js.level;
  out->ub = -1;
  out->lb = -1;
  const int H = q.D >= 0 ? q.D + 1 : 1;
  auto ext = [&](int32_t v) { return c->old_of_new.empty() ? v : c->old_of_new[v]; };
  const bool has_tree = js.ans != ~0ull && !js.err;
  if (!has_tree && own_edges && ebuf) std::free(ebuf);
  if (js.err == 7 && js.ans != ~0ull) {
    // the optimum is known; only its tree exceeds the edge buffer
    const long long idx = (long long)(js.ans & 0xffffffffull);
    out->status = GST_TREE_OMITTED;
    out->cost = out->ub = out->lb = (int32_t)(js.ans >> 32);
    out->root = ext((int32_t)(idx / H));
    out->root_height = q.D >= 0 ? (int32_t)(idx % H) : -1;
    return;
  }
  if (js.err) {
    out->status = GST_INTERNAL;
    out->cost = -1;
    char buf[64];
    std::snprintf(buf, sizeof buf, "tree reconstruction failed (code %d)", js.err);
    set_err(buf);
    return;
  }
  if (js.ans == ~0ull) {
    // saturated costs can hide a tree only where one exists in the graph at all
    const bool range = js.sat && comp_feasible(c, q);
    out->status = range ? GST_COST_RANGE : GST_NO_SOLUTION;
    out->cost = -1;
    if (range) set_err("costs reached the saturation range and no tree was found");
    return;
  }
  const long long idx = (long long)(js.ans & 0xffffffffull);
  out->status = GST_OK;
  out->cost = out->ub = out->lb = (int32_t)(js.ans >> 32);
  out->root = ext((int32_t)(idx / H));
  out->root_height = q.D >= 0 ? (int32_t)(idx % H) : -1;
  // distinct edges in the caller's ids, sorted
  struct E { int32_t a, b, w; };
  E *e = reinterpret_cast<E *>(ebuf);
  const int nr = ebuf ? n_raw : 0;
  for (int i = 0; i < nr; i++) {
    const int32_t a = ext(e[i].a), b = ext(e[i].b);
    e[i].a = std::min(a, b);
    e[i].b = std::max(a, b);
  }
  std::sort(e, e + nr, [](const E &x, const E &y) { return x.a != y.a ? x.a < y.a : x.b != y.b ? x.b < y.b : x.w < y.w; });
  int nu = 0;
  long long sum = 0;
  for (int i = 0; i < nr; i++) {
    if (nu && e[nu - 1].a == e[i].a && e[nu - 1].b == e[i].b) continue;
    e[nu++] = e[i];
    sum += e[i].w;
  }
  out->n_edges = nu;
  out->edges = (own_edges && nu) ? ebuf : nullptr;
  if (own_edges && ebuf && !nu) std::free(ebuf);
  out->tree_cost = (int32_t)sum;
}

}  // namespace

#include "gst_hubs.inc"
#include "gst_batch.inc"
#include "gst_cohort.inc"

namespace {
int noalloc_setup(gst_ctx *c);  // gst_noalloc.inc
}

namespace {

// The graph as the device holds it. With renumbering, device vertex i is the caller's
// vertex old_of_new[i], taken in order of descending degree (ties by id), and every group
// is re-sorted by device id. Every adjacency list is sorted by (weight, device id), with
// or without renumbering: the grow kernels stop scanning a row at the first arc whose
// weight lifts a state above the query's upper bound, since every later arc weighs at
// least as much (without renumbering only gverts points at the caller's array).
struct DeviceGraph {
  std::vector<int32_t> beg, old_of_new;
  std::vector<int32_t> adj_v, w_v, gverts_v;
  const int32_t *adj = nullptr, *w = nullptr, *gverts = nullptr;
};

template <class F>
void parallel_ranges(int64_t n, F f) {
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([=] { f(n * t / nt, n * (t + 1) / nt); });
  for (auto &x : th) x.join();
}

DeviceGraph renumber(const gst_csr *csr, const gst_groups *groups, bool reorder) {
  const int32_t n = csr->n;
  const int64_t m = csr->m;
  DeviceGraph d;
  d.beg.resize(n + 1);
  const int64_t total = groups->n_groups ? groups->offs[groups->n_groups] : 0;
  // row of (weight, id) keys, ascending
  auto sort_row = [](std::vector<uint64_t> &key, int32_t *adj_out, int32_t *w_out) {
    std::sort(key.begin(), key.end());
    for (size_t t = 0; t < key.size(); t++) {
      w_out[t] = (int32_t)(key[t] >> 32);
      adj_out[t] = (int32_t)(key[t] & 0xffffffffu);
    }
  };
  if (!reorder) {
    for (int32_t v = 0; v <= n; v++) d.beg[v] = (int32_t)csr->beg_pos[v];
    d.adj_v.resize(m);
    d.w_v.resize(m);
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
      std::vector<uint64_t> key;
      for (int64_t v = lo; v < hi; v++) {
        const int64_t b = csr->beg_pos[v], e = csr->beg_pos[v + 1];
        key.resize(e - b);
        for (int64_t t = b; t < e; t++) key[t - b] = ((uint64_t)(uint32_t)csr->w[t] << 32) | (uint32_t)csr->adj[t];
        sort_row(key, d.adj_v.data() + b, d.w_v.data() + b);
      }
    });
    d.gverts = groups->verts;
    return d;
  }
  d.old_of_new.resize(n);
  for (int32_t v = 0; v < n; v++) d.old_of_new[v] = v;
  auto deg = [&](int32_t v) { return csr->beg_pos[v + 1] - csr->beg_pos[v]; };
  std::stable_sort(d.old_of_new.begin(), d.old_of_new.end(),
                   [&](int32_t a, int32_t b) { return deg(a) > deg(b); });
  std::vector<int32_t> new_of_old(n);
  for (int32_t i = 0; i < n; i++) new_of_old[d.old_of_new[i]] = i;
  d.beg[0] = 0;
  for (int32_t i = 0; i < n; i++) d.beg[i + 1] = d.beg[i] + (int32_t)deg(d.old_of_new[i]);
  d.adj_v.resize(m);
  d.w_v.resize(m);
  parallel_ranges(n, [&](int64_t lo, int64_t hi) {
    std::vector<uint64_t> key;
    for (int64_t i = lo; i < hi; i++) {
      const int32_t o = d.old_of_new[i];
      const int64_t b = csr->beg_pos[o], e = csr->beg_pos[o + 1];
      key.resize(e - b);
      for (int64_t t = b; t < e; t++)
        key[t - b] = ((uint64_t)(uint32_t)csr->w[t] << 32) | (uint32_t)new_of_old[csr->adj[t]];
      sort_row(key, d.adj_v.data() + d.beg[i], d.w_v.data() + d.beg[i]);
    }
  });
  d.gverts_v.resize(total);
  for (int32_t g = 0; g < groups->n_groups; g++) {
    for (int64_t t = groups->offs[g]; t < groups->offs[g + 1]; t++) d.gverts_v[t] = new_of_old[groups->verts[t]];
    std::sort(d.gverts_v.begin() + groups->offs[g], d.gverts_v.begin() + groups->offs[g + 1]);
  }
  return d;
}

}  // namespace

extern "C" {

const char *gst_last_error(void) { return g_err; }

const char *gst_strerror(int code) {
  switch (code) {
    case GST_OK: return "ok";
    case GST_NO_SOLUTION: return "no feasible tree (a query group cannot be connected)";
    case GST_PENDING: return "pending";
    case GST_BAD_INPUT: return "bad input";
    case GST_OVER_BUDGET: return "DP arena does not fit the memory budget";
    case GST_DEVICE_ERROR: return "device error";
    case GST_INTERNAL: return "internal error";
    case GST_NO_DEVICE: return "no HIP device";
    case GST_TREE_OMITTED: return "optimum found; its tree exceeds GST_MAX_TREE_EDGES and is not returned";
    case GST_COST_RANGE: return "costs reached the int32 saturation range and no tree was found";
    default: return "unknown status";
  }
}

int64_t gst_arena_bytes(int32_t n, int32_t k, int32_t D) { return arena_bytes_for(n, k, D); }

int64_t gst_job_bytes(gst_ctx *c, int32_t k, int32_t D) {
  if (!c || k < 1 || k > MAXK || D > GST_MAX_D) return GST_BAD_INPUT;
  return (int64_t)slot_bytes(c, k, D < 0 ? -1 : D);
}

int gst_open(const gst_csr *csr, const gst_groups *groups, const gst_opts *opts, gst_ctx **out) {
  if (!csr || !groups || !out) { set_err("null argument"); return GST_BAD_INPUT; }
  *out = nullptr;
  const int32_t n = csr->n;
  const int64_t m = csr->m;
  // arc offsets are int32 on the device, and a kernel may form an offset up to one chunk
  // (at most 2^24 arcs) or one strip past the last arc: m stays 2^25 below 2^31
  if (n < 1 || m < 0 || m > GST_MAX_ARCS || !csr->beg_pos || (m && (!csr->adj || !csr->w))) {
    set_err("bad CSR sizes (need n >= 1, m <= 2^31 - 2^25)");
    return GST_BAD_INPUT;
  }
  if (csr->beg_pos[0] != 0 || csr->beg_pos[n] != m) { set_err("beg_pos must run from 0 to m"); return GST_BAD_INPUT; }
  for (int32_t v = 0; v < n; v++)
    if (csr->beg_pos[v + 1] < csr->beg_pos[v]) { set_err("beg_pos not monotone"); return GST_BAD_INPUT; }
  // arcs checked while the connected components are joined (union-find, path halving)
  std::vector<int32_t> comp(n);
  for (int32_t v = 0; v < n; v++) comp[v] = v;
  auto root = [&](int32_t x) {
    while (comp[x] != x) { comp[x] = comp[comp[x]]; x = comp[x]; }
    return x;
  };
  for (int32_t v = 0; v < n; v++) {
    for (int64_t e = csr->beg_pos[v]; e < csr->beg_pos[v + 1]; e++) {
      const int32_t u = csr->adj[e];
      if (u < 0 || u >= n) { set_err("adjacency id out of range"); return GST_BAD_INPUT; }
      if (csr->w[e] < 1 || csr->w[e] > (1 << 20)) { set_err("edge weight outside [1, 2^20]"); return GST_BAD_INPUT; }
      if (u > v) {
        const int32_t a = root(v), b = root(u);
        if (a != b) comp[a < b ? b : a] = a < b ? a : b;
      }
    }
  }
  for (int32_t v = 0; v < n; v++) comp[v] = root(v);
  if (groups->n_groups < 0 || (groups->n_groups > 0 && (!groups->offs || !groups->verts))) {
    set_err("bad group table"); return GST_BAD_INPUT;
  }
  const int64_t total = groups->n_groups ? groups->offs[groups->n_groups] : 0;
  if (groups->n_groups && groups->offs[0] != 0) { set_err("group offsets must start at 0"); return GST_BAD_INPUT; }
  for (int32_t gI = 0; gI < groups->n_groups; gI++) {
    if (groups->offs[gI + 1] < groups->offs[gI]) { set_err("group offsets not monotone"); return GST_BAD_INPUT; }
    for (int64_t t = groups->offs[gI]; t < groups->offs[gI + 1]; t++) {
      if (groups->verts[t] < 0 || groups->verts[t] >= n) { set_err("group member out of range"); return GST_BAD_INPUT; }
      if (t > groups->offs[gI] && groups->verts[t] < groups->verts[t - 1]) { set_err("group members must be ascending"); return GST_BAD_INPUT; }
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    set_err("no HIP device");
    return GST_NO_DEVICE;
  }
  auto c = std::make_unique<gst_ctx>();
  if (opts) c->opts = *opts;
  c->dev = c->opts.device;
  if (c->dev < 0 || c->dev >= ndev) { set_err("bad device ordinal"); return GST_BAD_INPUT; }
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  {
    // the merges walk their tiles with a grid-stride loop: a grid of exactly the
    // resident blocks leaves no partial second wave of blocks behind the first
    int ncu = 0, nb = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_merge_hop<false>, BLK, 0));
    c->merge_hop_grid = (unsigned)std::max(1, ncu * nb);
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_merge_hop<true>, BLK, 0));
    c->merge_hopb_grid = (unsigned)std::max(1, ncu * nb);
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_merge, BLK, 0));
    c->merge_grid = (unsigned)std::max(1, ncu * nb);
  }
  c->n = n;
  c->m = m;
  c->n_groups = groups->n_groups;
  const int32_t split0 = c->opts.degree_split > 0 ? c->opts.degree_split : 32;
  const int32_t chunk0 = c->opts.chunk_edges > 0 ? ((c->opts.chunk_edges + 63) / 64) * 64 : 1024;
  c->delta = c->opts.near_far_delta > 0 ? c->opts.near_far_delta : 0;
  c->max_batch = c->opts.max_batch > 0 ? c->opts.max_batch : 64;
  DeviceGraph dg = renumber(csr, groups, !(c->opts.flags & GST_O_KEEP_ORDER));
  const std::vector<int32_t> &beg32 = dg.beg;
  c->h_deg.resize(n);
  for (int32_t v = 0; v < n; v++) c->h_deg[v] = beg32[v + 1] - beg32[v];
  c->old_of_new = std::move(dg.old_of_new);
  c->h_goffs.assign(groups->offs, groups->offs + groups->n_groups + 1);
  if (groups->n_groups == 0) c->h_goffs.assign(1, 0);
  c->h_gverts.assign(groups->verts, groups->verts + total);
  c->comp = std::move(comp);
  HIPCHK(hipMalloc((void **)&c->d_beg, sizeof(int32_t) * (n + 1)));
  HIPCHK(hipMalloc((void **)&c->d_adj, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  HIPCHK(hipMalloc((void **)&c->d_w, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  HIPCHK(hipMalloc((void **)&c->d_goffs, sizeof(long long) * c->h_goffs.size()));
  HIPCHK(hipMalloc((void **)&c->d_gverts, sizeof(int32_t) * std::max<int64_t>(total, 1)));
  HIPCHK(hipMemcpy(c->d_beg, beg32.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
  if (m) {
    HIPCHK(hipMemcpy(c->d_adj, dg.adj ? dg.adj : dg.adj_v.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_w, dg.w ? dg.w : dg.w_v.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
  }
  {
    std::vector<long long> go(c->h_goffs.begin(), c->h_goffs.end());
    HIPCHK(hipMemcpy(c->d_goffs, go.data(), sizeof(long long) * go.size(), hipMemcpyHostToDevice));
  }
  if (total) HIPCHK(hipMemcpy(c->d_gverts, dg.gverts ? dg.gverts : dg.gverts_v.data(), sizeof(int32_t) * total,
                              hipMemcpyHostToDevice));
  c->graph_bytes = (int64_t)4 * (n + 1) + 8 * m + 8 * (int64_t)c->h_goffs.size() + 4 * total;
  HIPCHK(hipMalloc((void **)&c->d_jc, sizeof(JobConst) * c->max_batch));
  HIPCHK(hipMalloc((void **)&c->d_js, sizeof(JobState) * c->max_batch));
  HIPCHK(hipHostMalloc((void **)&c->h_js, sizeof(JobState) * c->max_batch, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&c->h_jc, sizeof(JobConst), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&c->h_stx, sizeof(unsigned long long) * NSTRIPE * NSTX, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&c->h_count, sizeof(unsigned long long), hipHostMallocDefault));
  HIPCHK(hipMalloc((void **)&c->d_count, sizeof(unsigned long long)));
  HIPCHK(hipMalloc((void **)&c->d_plan, sizeof(Plan)));
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  HIPCHK(hipEventCreate(&c->cev0));
  HIPCHK(hipEventCreate(&c->cev1));
  c->g.beg = c->d_beg; c->g.adj = c->d_adj; c->g.w = c->d_w;
  c->g.goffs = c->d_goffs; c->g.gverts = c->d_gverts; c->g.n = n;
  {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    c->plain_cap = std::min<int64_t>((int64_t)fr / 8, (int64_t)32 << 30);
  }
  int rc = gst_set_tuning(c.get(), split0, chunk0, c->delta);
  if (rc) return rc;
  if (!(c->opts.flags & GST_O_NO_HUB_INDEX)) {
    rc = hub_index_build(c.get());
    if (rc) return rc;
  }
  rc = batch_open(c.get());
  if (rc) return rc;
  rc = noalloc_setup(c.get());
  if (rc) return rc;
  *out = c.release();
  return GST_OK;
}

}  // extern "C"

namespace {

// Index of the tuning (split, chunk) in c->tunes, building its chunk counts when new.
// A new tuning whose wavefront items exceed the worklist capacity rebuilds the slots.
int tune_index(gst_ctx *c, int32_t split, int32_t chunk, int &idx) {
  for (size_t i = 0; i < c->tunes.size(); i++)
    if (c->tunes[i].split == split && c->tunes[i].chunk == chunk) { idx = (int)i; return GST_OK; }
  // drop tunings no k uses any more (each holds n chunk counts on the device)
  for (size_t i = c->tunes.size(); i-- > 0;) {
    bool used = false;
    for (int k = 0; k <= MAXK; k++) used = used || c->tune_k[k] == (int)i;
    if (used) continue;
    (void)hipFree(c->tunes[i].d_nchunk);
    c->tunes.erase(c->tunes.begin() + i);
    for (int k = 0; k <= MAXK; k++)
      if (c->tune_k[k] > (int)i) c->tune_k[k]--;
  }
  gst_ctx::TuneSet ts;
  ts.split = split;
  ts.chunk = chunk;
  std::vector<int32_t> nch(c->n, 0);
  for (int32_t v = 0; v < c->n; v++)
    if (c->h_deg[v] >= split) {
      nch[v] = (c->h_deg[v] + chunk - 1) / chunk;
      ts.items += nch[v];
    }
  HIPCHK(hipMalloc((void **)&ts.d_nchunk, sizeof(int32_t) * std::max<int32_t>(c->n, 1)));
  HIPCHK(hipMemcpy(ts.d_nchunk, nch.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
  // with vertices renumbered by descending degree the wavefront vertices are a prefix
  ts.n_heavy = c->n;
  if (!c->old_of_new.empty()) {
    int32_t h = 0;
    while (h < c->n && c->h_deg[h] >= split) h++;
    bool prefix = true;
    for (int32_t v = h; v < c->n && prefix; v++) prefix = c->h_deg[v] < split;
    if (prefix) ts.n_heavy = h;
  }
  const long long bcap = std::max<long long>(ts.items, 1);
  if (bcap > c->bcap) {
    // slots sized for the old capacity must be rebuilt
    for (auto &e : c->plain) slot_free(e.second);
    c->plain.clear();
    batch_release_slots(c);
    c->bcap = bcap;
  }
  c->tunes.push_back(ts);
  idx = (int)c->tunes.size() - 1;
  return GST_OK;
}

int check_tuning(int32_t degree_split, int32_t chunk_edges) {
  // a wavefront of thread items sums 64 degrees below the split in 32 bits
  if (degree_split < 1 || degree_split > (1 << 24) || chunk_edges < 64 || chunk_edges > (1 << 24)) {
    set_err("bad tuning (need 1 <= degree_split <= 2^24, 64 <= chunk_edges <= 2^24)");
    return GST_BAD_INPUT;
  }
  return GST_OK;
}

}  // namespace

extern "C" {

int gst_set_tuning(gst_ctx *c, int32_t degree_split, int32_t chunk_edges, int32_t near_far_delta) {
  if (!c || near_far_delta < 0) { set_err("bad tuning (need near_far_delta >= 0)"); return GST_BAD_INPUT; }
  if (int rc = check_tuning(degree_split, chunk_edges)) return rc;
  chunk_edges = ((chunk_edges + 63) / 64) * 64;
  for (int k = 0; k <= MAXK; k++) c->tune_k[k] = -1;  // every tuning becomes unused ...
  int idx = 0;
  if (int rc = tune_index(c, degree_split, chunk_edges, idx)) return rc;
  for (int k = 0; k <= MAXK; k++) { c->tune_k[k] = idx; c->tile_k[k] = 0; }  // ... but this one
  c->delta = near_far_delta;
  return GST_OK;
}

int gst_set_tuning_k(gst_ctx *c, int32_t k, int32_t degree_split, int32_t chunk_edges, int32_t merge_tile) {
  if (!c || k < 1 || k > MAXK || merge_tile < 0) { set_err("bad tuning (need 1 <= k <= GST_MAX_K, merge_tile >= 0)"); return GST_BAD_INPUT; }
  if (int rc = check_tuning(degree_split, chunk_edges)) return rc;
  chunk_edges = ((chunk_edges + 63) / 64) * 64;
  const int old = c->tune_k[k];
  c->tune_k[k] = -1;
  int idx = 0;
  const int rc = tune_index(c, degree_split, chunk_edges, idx);
  if (rc) { c->tune_k[k] = old; return rc; }
  c->tune_k[k] = idx;
  c->tile_k[k] = merge_tile;
  return GST_OK;
}

static int solve_on_slot(gst_ctx *c, Slot &s, const Query &q, int32_t *ebuf, int32_t ecap_caller, gst_result *out);

int gst_solve(gst_ctx *c, const int32_t *gids, int32_t k, int32_t D, int32_t flags, gst_result *out) {
  if (!c || !out) { set_err("null argument"); return GST_BAD_INPUT; }
  std::memset(out, 0, sizeof(*out));
  // until the result is complete it reads as a device error (every early return below
  // that is not a query error is one)
  out->status = GST_DEVICE_ERROR;
  out->cost = -1;
  out->ub = out->lb = -1;
  int rc;
  {
    std::lock_guard<std::mutex> lk(c->solve_mu);
    Query q;
    rc = check_query(c, gids, k, D, flags, q);
    if (rc == GST_OK) rc = hipSetDevice(c->dev) == hipSuccess ? plain_slot(c, q.K, q.D) : GST_DEVICE_ERROR;
    if (rc == GST_OK) rc = solve_on_slot(c, c->plain.front().second, q, nullptr, 0, out);
  }
  if (rc < 0 && rc != GST_INTERNAL && rc != GST_COST_RANGE) {
    gst_result_free(out);
    out->n_edges = 0;
    out->status = rc;
    out->cost = -1;
  }
  return rc;
}

// One query on slot s, host-driven rounds. With ebuf == nullptr the tree goes to a fresh
// heap buffer owned by the result; otherwise into ebuf (ecap_caller triples), with no
// host or device allocation on the way.
static int solve_on_slot(gst_ctx *c, Slot &s, const Query &q, int32_t *ebuf, int32_t ecap_caller, gst_result *out) {
  const auto t0 = clk::now();
  JobConst &jc = *c->h_jc;
  fill_jobconst(c, s, q, jc);
  JobState &js = c->h_js[0];
  init_job(c, q, js);
  HIPCHK(hipMemcpyAsync(c->d_jc, &jc, sizeof(jc), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_js, &js, sizeof(js), hipMemcpyHostToDevice, c->st));
  launch_job_init(c->st, s, c, jc);
  HIPCHK(hipGetLastError());
  const GraphDev g = c->g;
  JobState &h = c->h_js[0];
  double t_dp = 0;
  int guard = 0;
  bool dp_round = false;  // previous round was part of the DP loop
  int timed_phase = -1;   // phase whose work kernels ev0/ev1 bracket
  double t_grow = 0, t_merge = 0;
  auto tr0 = clk::now();
  for (;;) {
    k_step<<<1, BLK, 0, c->st>>>(c->d_jc, c->d_js, c->d_plan, 1, c->n);
    HIPCHK(hipMemcpyAsync(&h, c->d_js, sizeof(JobState), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (dp_round) t_dp += secs(tr0, clk::now());
    if (timed_phase == PH_GROW || timed_phase == PH_MERGE) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      (timed_phase == PH_GROW ? t_grow : t_merge) += ms * 1e-3;
    }
    timed_phase = -1;
    tr0 = clk::now();
    if (h.phase == PH_DONE) break;
    if (++guard > (1 << 22)) { set_err("round limit exceeded"); out->status = GST_INTERNAL; return GST_INTERNAL; }
    const int ph = h.phase;
    Plan *pl = c->d_plan;
    switch (ph) {
      case PH_SEED:
        k_seed<<<grid_for(jc.seed_total, BLK, 1024), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      case PH_GROW: {
        const int li = q.D >= 0 ? h.depth : h.parity;
        const int n0 = h.cnt[li][0], n1 = h.cnt[li][1];
        HIPCHK(hipEventRecord(c->ev0, c->st));
        k_select<<<grid_for(n0 + n1, BLK, GROW_GRID_CAP), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        if (n0) k_grow_thread<<<grid_for(n0, BLK, GROW_GRID_CAP), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        if (n1) k_grow_wave<<<grid_for(n1, BLK / WAVE, GROW_GRID_CAP), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        HIPCHK(hipEventRecord(c->ev1, c->st));
        timed_phase = PH_GROW;
        k_clear<<<grid_for(n0 + n1, BLK, 1 << 16), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      }
      case PH_FARSPLIT:
        k_farsplit<<<grid_for(h.cnt[FAR_LIST + h.fpar][0], BLK, 1 << 16), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      case PH_HOPCLEAR: {
        const int n0 = h.cnt[q.D][0], n1 = h.cnt[q.D][1];
        k_clear<<<grid_for(n0 + n1, BLK, 1 << 16), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      }
      case PH_TABLES:
        k_tables<<<1, BLK, 0, c->st>>>(c->d_jc, c->d_js, 1, g);
        if (jc.sparse) k_alive<<<grid_for(c->n, BLK, GROW_GRID_CAP), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      case PH_MERGE: {
        const bool bf = hop_bound_first(jc, h.level);
        const int tile = bf ? jc.tvb : jc.tv;
        const long long tiles = ((long long)c->n + tile - 1) / tile;
        HIPCHK(hipEventRecord(c->ev0, c->st));
        if (bf) k_merge_hop<true><<<grid_for(tiles, 1, c->merge_hopb_grid), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        else if (q.D >= 0) k_merge_hop<false><<<grid_for(tiles, 1, c->merge_hop_grid), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        else k_merge<<<grid_for(tiles, 1, c->merge_grid), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        if (h.level < q.K && (q.D != 0)) {
          const long long nblk = ((long long)c->n + 63) / 64 * (q.D > 0 ? q.D : 1);
          k_collect<<<grid_for(nblk, BLK / WAVE, GROW_GRID_CAP), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        }
        HIPCHK(hipEventRecord(c->ev1, c->st));
        timed_phase = PH_MERGE;
        break;
      }
      case PH_FINAL:
        k_final<<<grid_for(((long long)c->n * jc.H + FINAL_CHUNK - 1) / FINAL_CHUNK, 1, 4096), BLK, 0, c->st>>>(c->d_jc, c->d_js, pl, g);
        break;
      case PH_REBUILD:
        k_rebuild<<<1, RB_BLK, 0, c->st>>>(c->d_jc, c->d_js, 1, g);
        break;
      default:
        break;
    }
    HIPCHK(hipGetLastError());
    // DP-loop time: merge and grow rounds after the level-1 SSSP (hop: every round)
    dp_round = (ph == PH_MERGE || ph == PH_GROW || ph == PH_HOPCLEAR) && (q.D >= 0 || h.level >= 2);
  }
  const int n_out = std::min(std::max(h.n_out, 0), s.ecap);
  int32_t *eb = nullptr;
  if (n_out > 0 && h.ans != ~0ull && !h.err) {
    eb = ebuf ? ebuf : (int32_t *)std::malloc(sizeof(int32_t) * 3 * n_out);
    if (ebuf && n_out > ecap_caller) { set_err("tree exceeds edge_cap"); return GST_INTERNAL; }
    const hipError_t e = hipMemcpy(eb, s.out, sizeof(int32_t) * 3 * n_out, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      if (!ebuf) std::free(eb);
      std::snprintf(g_err, sizeof g_err, "tree copy: %s", hipGetErrorString(e));
      return GST_DEVICE_ERROR;
    }
  }
  JobState fin = h;
  HIPCHK(hipMemcpy(c->h_stx, s.stx, (size_t)NSTRIPE * NSTX * 8, hipMemcpyDeviceToHost));
  fold_stats(c->h_stx, fin);
  finish_result(c, q, fin, eb, n_out, ebuf == nullptr, out);
  if (q.flags & GST_F_STATS) {
    k_fill_u64<<<1, 64, 0, c->st>>>(c->d_count, 1, 0ull);
    // the states of levels 1..k (the arena's last n * H words are unused and not filled)
    const long long used = std::min<long long>((long long)s.arena_words, jc.off[q.K + 1]);
    k_count_finite<<<grid_for(used, BLK, 4096), BLK, 0, c->st>>>(s.T, used, c->d_count);
    HIPCHK(hipMemcpyAsync(c->h_count, c->d_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    out->states_finite = (int64_t)*c->h_count;
  }
  out->t_dp_s = t_dp;
  out->t_grow_s = t_grow;
  out->t_merge_s = t_merge;
  out->t_total_s = secs(t0, clk::now());
  return out->status < 0 ? out->status : GST_OK;
}

void gst_result_free(gst_result *r) {
  if (r && r->edges) { std::free(r->edges); r->edges = nullptr; }
}

int gst_info_get(gst_ctx *c, gst_info *info) {
  if (!c || !info) return GST_BAD_INPUT;
  std::memset(info, 0, sizeof(*info));
  info->n = c->n;
  info->m = c->m;
  info->graph_bytes = c->graph_bytes;
  batch_info(c, info);
  return GST_OK;
}

int gst_get_cohort_stats(gst_ctx *c, gst_cohort_stats *out, int32_t reset) {
  if (!c || !out) return GST_BAD_INPUT;
  std::lock_guard<std::mutex> lk(c->cstat_mu);
  *out = c->cstat;
  if (reset) c->cstat = gst_cohort_stats{};
  return GST_OK;
}

void gst_close(gst_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  batch_close(c);
  for (auto &e : c->plain) slot_free(e.second);
  c->plain.clear();
  (void)hipFree(c->d_beg); (void)hipFree(c->d_adj); (void)hipFree(c->d_w);
  (void)hipFree(c->d_goffs); (void)hipFree(c->d_gverts);
  for (auto &t : c->tunes) (void)hipFree(t.d_nchunk);
  (void)hipFree(c->d_jc); (void)hipFree(c->d_js); (void)hipHostFree(c->h_js);
  (void)hipHostFree(c->h_jc); (void)hipHostFree(c->h_stx); (void)hipHostFree(c->h_count);
  (void)hipFree(c->d_count);
  (void)hipFree(c->d_plan);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->cev0) (void)hipEventDestroy(c->cev0);
  if (c->cev1) (void)hipEventDestroy(c->cev1);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

}  // extern "C"

#include "gst_noalloc.inc"
