// gst_cohort_kernels.hpp — device code of cohorts: up to 64 queries of one k whose grow
// phases run together (gfx950, wave64).
//
// A cohort's grow rounds relax one level block in which the row of vertex v holds the
// states of every member at that level, member after member:
//   CT[v * S + q * C + j] = T_q(v, j-th subset of size L),  C = C(k, L), S = cohort_stride.
// A relaxation of arc (v, u) then reads u's whole row with coalesced loads, 64 labels per
// wavefront load, instead of one random cache line per member: on a power-law graph the
// members' frontiers overlap almost completely (most vertices are active in the same
// rounds for most labels), so one adjacency scan and one row fetch serve all of them.
// The merges, tour tables, final minimum and tree rebuild stay per member (the per-query
// kernels, on "passive" jobs whose phases the host sets): a member's merge writes its
// level into the cohort block directly, and the block is copied back into the members'
// own level-major arenas after the level's grow rounds (k_cscatter), where the later
// merges, the lower bounds and the rebuild read it.
#pragma once
#include "gst_kernels.hpp"

namespace gst {

constexpr int CMAXB = 64;   // members per cohort
constexpr int CMAXCH = 16;  // 64-label chunks per cohort row (row width <= 1024 labels)

struct CohortConst {
  int32_t *CT;               // [n][S] states of the current level
  unsigned long long *cact;  // [2][n][NW] label masks per list (bit l = label l of the row)
  int32_t *flag;             // [2][n] vertex activated in this list (next frontier)
  int32_t *items;            // [2][icap][2] (vertex, chunk) frontier items; chunk -1 = all arcs
  unsigned long long *stx;   // [NSTRIPE][NSTX] cohort work counters
  long long icap;
  const int32_t *nchunk;     // the members' grow tuning (one k, so one tuning)
  int32_t chunk, n_heavy;
  int32_t B, K, L, C, S, NW; // members, k, level, labels per member, row words, chunks
  int32_t jobs[CMAXB];       // member index -> job index in the cohort's job table
};

struct CohortState {
  int32_t par;     // list consumed this round
  int32_t cnt[2];  // items per list
};

// Bits [l0, l0 + C) of a row's label masks (C <= 64).
__device__ __forceinline__ unsigned long long mask_bits(const unsigned long long *row, int l0, int C) {
  const int w0 = l0 >> 6, sh = l0 & 63;
  unsigned long long a = row[w0] >> sh;
  if (sh + C > 64) a |= row[w0 + 1] << (64 - sh);
  return C == 64 ? a : a & ((1ull << C) - 1);
}

// Seeds: T_q(x, {i}) = 0 for every member x of member q's group i (one block per member).
__global__ __launch_bounds__(BLK) void k_cseed(const CohortConst *__restrict__ ccp, const JobConst *__restrict__ jcs,
                                               GraphDev g) {
  const CohortConst &cc = *ccp;
  const int q = blockIdx.x;
  if (q >= cc.B) return;
  const JobConst &c = jcs[cc.jobs[q]];
  for (int i = 0; i < c.K; i++) {
    const long long b0 = g.goffs[c.gid[i]], b1 = g.goffs[c.gid[i] + 1];
    const int l = q * cc.C + i;
    for (long long t = b0 + threadIdx.x; t < b1; t += BLK) {
      const int32_t x = g.gverts[t];
      cc.CT[(size_t)x * cc.S + l] = 0;
      atomicOr(&cc.cact[((size_t)1 * g.n + x) * cc.NW + (l >> 6)], 1ull << (l & 63));
      cc.flag[(size_t)1 * g.n + x] = 1;
    }
  }
}

// Frontier of list nx: every flagged vertex becomes one item (or one per chunk of arcs
// for vertices of degree >= split); the flag is cleared. Wavefront-aggregated appends.
__global__ __launch_bounds__(BLK) void k_ccollect(const CohortConst *__restrict__ ccp, CohortState *cs, GraphDev g,
                                                  int nx) {
  const CohortConst &cc = *ccp;
  gi32 *flag = as_global(cc.flag + (size_t)nx * g.n);
  int32_t *items = cc.items + (size_t)nx * cc.icap * 2;
  const long long stride = (long long)gridDim.x * BLK;
  for (long long b = (long long)blockIdx.x * BLK + (threadIdx.x & ~63); b < g.n; b += stride) {
    const long long v = b + lane_id();
    int cnt = 0, nch = 0;
    if (v < g.n && flag[v]) {
      flag[v] = 0;
      nch = chunks_of(cc.nchunk, cc.n_heavy, (int32_t)v);
      cnt = nch ? nch : 1;
    }
    const int pos = wave_append(&cs->cnt[nx], cnt);
    if (cnt) {
      gi32 *it = as_global(items) + (size_t)pos * 2;
      if (!nch) { it[0] = (int32_t)v; it[1] = -1; }
      for (int i = 0; i < nch; i++) { it[2 * i] = (int32_t)v; it[2 * i + 1] = i; }
    }
  }
}

// Select pass: a wavefront per frontier vertex, lane q = member q, which evaluates its
// active states at v as the per-query select pass does (UB candidates, UB/2 and lower
// bound cuts) and drops the bits of the states that do not grow.
__global__ __launch_bounds__(BLK) void k_cselect(const CohortConst *__restrict__ ccp, const CohortState *__restrict__ csp,
                                                 const JobConst *jcs, JobState *jss, GraphDev g) {
  const CohortConst &cc = *ccp;
  const int cur = csp->par;
  const long long total = csp->cnt[cur];
  const int lane = lane_id();
  const long long waves = (long long)gridDim.x * (BLK / WAVE);
  const int L = cc.L, C = cc.C;
  unsigned long long n_states = 0;
  for (long long it = (long long)blockIdx.x * (BLK / WAVE) + (threadIdx.x >> 6); it < total; it += waves) {
    const int32_t *item = cc.items + ((size_t)cur * cc.icap + it) * 2;
    if (item[1] > 0) continue;  // one evaluation per vertex
    const int32_t v = item[0];
    if (lane >= cc.B) continue;
    unsigned long long *row_m = cc.cact + ((size_t)cur * g.n + v) * cc.NW;
    const int l0 = lane * C;
    const unsigned long long a = mask_bits(row_m, l0, C);
    if (!a) continue;
    const JobConst &c = jcs[cc.jobs[lane]];
    JobState *s = &jss[cc.jobs[lane]];
    const int32_t *row = cc.CT + (size_t)v * cc.S + l0;
    const int32_t *sing = L == 1 ? row : lrow(c, 1, v, 0);
    int ns = 0, dmin = INF;
    unsigned long long defer = 0;
    const unsigned long long gm = grow_select_rows(c, s, v, row, sing, a, L, load_ub(s), ns, INF, defer, dmin);
    n_states += ns;
    const unsigned long long drop = a & ~gm;
    if (drop) {
      const int w0 = l0 >> 6, sh = l0 & 63;
      atomicAnd(&row_m[w0], ~(drop << sh));
      if (sh + C > 64) atomicAnd(&row_m[w0 + 1], ~(drop >> (64 - sh)));
    }
  }
  // lane q counted member q's evaluated states over this wave's items
  if (lane < cc.B && n_states)
    atomicAdd(jcs[cc.jobs[lane]].stx + (blockIdx.x & (NSTRIPE - 1)) * NSTX + ST_STATES, n_states);
}

// Grow: a wavefront per item (vertex, chunk of arcs); lanes over the row's labels, 64 per
// chunk, NWT chunks at most (the row's chunk count rounded up to an instantiated one, so
// that the per-arc arrays stay in registers). v's values of its active labels stay in
// registers; for each arc (u, w), the target row's active chunks are read with coalesced
// loads, AU arcs in flight at a time; an improvement is an atomicMin (fire-and-forget), the improved bits are OR-ed
// into u's masks of the next list by one lane per chunk, and u is flagged for the next
// frontier (k_ccollect builds it).
// The UB/2 rule is applied here again, with each member's UB as it stands when the item
// starts (it only decreases after the select pass): a source above floor(UB/2) is not
// grown (in an optimal tree rooted at its weighted centroid every grow step starts from a
// subtree of weight <= OPT/2). Targets up to the member's UB are stored (a grow result
// above OPT/2 can still be a merge input at the centroid); a relaxation that would lift a
// label above its member's UB is not made, and since rows are sorted by weight the arc
// scan of an item ends at the first arc too heavy for every grown label.
template <int NWT, int AU>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(NWT <= 5 ? 5 : 1))) void k_cgrow(const CohortConst *__restrict__ ccp, const CohortState *__restrict__ csp,
                                               const JobConst *jcs, const JobState *jss, GraphDev g) {
  const CohortConst &cc = *ccp;
  const int cur = csp->par, nx = cur ^ 1;
  const long long total = csp->cnt[cur];
  const int lane = lane_id();
  const int S = cc.S, NW = cc.NW;
  const long long waves = (long long)gridDim.x * (BLK / WAVE);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  gi32 *CT = as_global(cc.CT);
  const gu64 *actc = as_global(cc.cact + (size_t)cur * g.n * NW);
  gu64 *actn = as_global(cc.cact + (size_t)nx * g.n * NW);
  gi32 *flagn = as_global(cc.flag + (size_t)nx * g.n);
  const gi32 *itm = as_global(cc.items + (size_t)cur * cc.icap * 2);
  const int32_t *__restrict__ adj = g.adj;
  const int32_t *__restrict__ wts = g.w;
  // per label of a row: its member's UB (INF when the member does not prune), read once
  // per block instead of through two dependent loads per item. UB only decreases, so the
  // value at launch bounds every optimal derivation through the launch (at level 1 it
  // does not change during the rounds at all: k_cstar lowers it after them).
  __shared__ int32_t lub[NWT * 64];
  for (int l = threadIdx.x; l < NWT * 64; l += BLK) {
    int32_t x = INF;
    const int q = l / cc.C;
    if (l < S && q < cc.B) {
      const int jb = cc.jobs[q];
      if (!(jcs[jb].flags & JF_NO_PRUNE)) x = load_ub(&jss[jb]);
    }
    lub[l] = x;
  }
  __syncthreads();
  unsigned long long n_edges = 0, n_relax = 0;
  for (long long it = (long long)blockIdx.x * (BLK / WAVE) + wave; it < total; it += waves) {
    const int32_t v = __builtin_amdgcn_readfirstlane(itm[2 * it]);
    const int32_t ch = __builtin_amdgcn_readfirstlane(itm[2 * it + 1]);
    int32_t sv[NWT], cap[NWT];  // cap: the member's growth cap, the largest value activated
    int32_t lim[NWT];           // the heaviest arc a grown label relaxes: UB - its value
    int nb = 0;
#pragma unroll
    for (int c = 0; c < NWT; c++) {
      unsigned long long x = c < NW ? actc[(size_t)v * NW + c] : 0ull;
      x = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(x >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
      const bool bit = (x >> lane) & 1;
      sv[c] = bit ? CT[(size_t)v * S + c * 64 + lane] : INF;
      const int ubq = lub[c * 64 + lane];
      cap[c] = ubq < INF ? grow_cap(ubq, cc.L) : INF;
      const bool grow = bit && sv[c] <= cap[c];
      lim[c] = grow ? ubq - sv[c] : -1;
      nb += __popcll(__ballot(grow));
    }
    if (!nb) continue;
    // rows are sorted by weight: the scan ends at the first arc heavier than every lim,
    // and a chunk is skipped for arcs heavier than all of its lims (cmx: per chunk)
    int32_t cmx[NWT];
#pragma unroll
    for (int c = 0; c < NWT; c++) {
      int x = lim[c];
      for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
      }
      cmx[c] = __builtin_amdgcn_readfirstlane(x);
    }
    int lmax = -1;
#pragma unroll
    for (int c = 0; c < NWT; c++) lmax = cmx[c] > lmax ? cmx[c] : lmax;
    const int32_t vb = g.beg[v], ve = g.beg[v + 1];
    const int32_t eb = ch < 0 ? vb : vb + ch * cc.chunk;
    const int32_t ee = ch < 0 ? ve : min(ve, eb + cc.chunk);
    uint32_t nr = 0;  // relaxations of this lane in this item
    for (int32_t e0 = eb; e0 < ee; e0 += WAVE) {
      const int na = min(WAVE, ee - e0);
      const int32_t au = lane < na ? adj[e0 + lane] : 0;
      const int32_t aw = lane < na ? wts[e0 + lane] : 0;
      if (__builtin_amdgcn_readfirstlane(aw) > lmax) break;  // lane 0 holds arc e0
      n_edges += na;
      for (int j = 0; j < na; j += AU) {  // uniform trip count
        // wt: the arc's weight, or 2^31 - 1 past the strip's arcs, so that one uniform
        // test against a chunk's lim maximum also covers the strip's end
        int32_t u[AU], wt[AU];
#pragma unroll
        for (int a = 0; a < AU; a++) {
          u[a] = __builtin_amdgcn_readlane(au, min(j + a, WAVE - 1));
          wt[a] = j + a < na ? __builtin_amdgcn_readlane(aw, min(j + a, WAVE - 1)) : 0x7fffffff;
        }
        // t: the target label, 0 where no relaxation is made (no candidate, >= 1, is
        // below it); lanes of labels not grown have lim = -1 and make none
        int32_t t[AU][NWT];
#pragma unroll
        for (int a = 0; a < AU; a++)
#pragma unroll
          for (int c = 0; c < NWT; c++) {
            t[a][c] = 0;
            if (wt[a] <= cmx[c]) {  // uniform
              const bool r = wt[a] <= lim[c];
              nr += (uint32_t)r;
              if (r) t[a][c] = CT[(size_t)u[a] * S + c * 64 + lane];
            }
          }
#pragma unroll
        for (int a = 0; a < AU; a++) {
          if (wt[a] > lmax) continue;  // uniform
          // lane c: the improved bits of chunk c to activate
          int32_t imlo = 0, imhi = 0;
          bool any = false;
#pragma unroll
          for (int c = 0; c < NWT; c++) {
            if (wt[a] > cmx[c]) continue;  // uniform (also every chunk without grown labels)
            int nc = sv[c] + wt[a];
            nc = nc < INF ? nc : INF;
            const bool imp = nc < t[a][c];
            // activated only where the select pass would grow it (<= the member's cap)
            const bool act = imp && nc <= cap[c];
            if (imp)
              __hip_atomic_fetch_min(&CT[(size_t)u[a] * S + c * 64 + lane], nc, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long im = __builtin_amdgcn_ballot_w64(act);
            if (im) {  // uniform
              if (lane == c) {
                imlo = (int32_t)(uint32_t)im;
                imhi = (int32_t)(uint32_t)(im >> 32);
              }
              any = true;
            }
          }
          if (any) {
            const unsigned long long imc = ((unsigned long long)(uint32_t)imhi << 32) | (uint32_t)imlo;
            if (imc)
              __hip_atomic_fetch_or(&actn[(size_t)u[a] * NW + lane], imc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lane == 0) flagn[u[a]] = 1;
          }
        }
      }
    }
    n_relax += nr;
  }
  for (int off = 32; off > 0; off >>= 1) n_relax += __shfl_xor(n_relax, off, 64);  // per-lane counts
  if (lane == 0 && n_edges) {
    unsigned long long *st = cc.stx + (blockIdx.x & (NSTRIPE - 1)) * NSTX;
    atomicAdd(st + ST_EDGES, n_edges);
    atomicAdd(st + ST_RELAX, n_relax);
    atomicAdd(st + SX_LEVEL + 2 * (cc.L & 7), n_edges);
    atomicAdd(st + SX_LEVEL + 2 * (cc.L & 7) + 1, n_relax);
  }
}

// Level-1 stars of every member over every vertex, after the cohort's level-1 rounds (the
// UB candidates a level-1 select pass would offer): a wavefront per vertex, lane q =
// member q, sum_i T_q(v, {i}) when every label is finite; each lane keeps its minimum and
// lowers its member's UB once at the end. A label within reach of INF flags its member's
// saturation check, as the select pass does. Each member's finite level-1 labels are
// counted into its evaluated-state counter (level 1 has no select pass to count them: a
// state reached at level 1 counts once, as a best-first search takes it once).
__global__ __launch_bounds__(BLK) void k_cstar(const CohortConst *__restrict__ ccp, const JobConst *jcs, JobState *jss,
                                               GraphDev g) {
  const CohortConst &cc = *ccp;
  const int lane = lane_id();
  long long best = INF;
  bool sat = false;
  unsigned long long reached = 0;
  if (lane < cc.B) {
    const gi32 *CT = as_global(cc.CT);
    for (long long v = (long long)blockIdx.x * (BLK / WAVE) + (threadIdx.x >> 6); v < g.n;
         v += (long long)gridDim.x * (BLK / WAVE)) {
      long long sum = 0;
      for (int i = 0; i < cc.K; i++) {
        const int x = CT[(size_t)v * cc.S + lane * cc.C + i];
        sat = sat || (x >= SAT_AT && x < INF);
        reached += x < INF;
        sum = x >= INF ? (long long)INF : sum + x;
      }
      best = sum < best ? sum : best;
    }
    JobState *s = &jss[cc.jobs[lane]];
    if (best < INF && !(jcs[cc.jobs[lane]].flags & JF_NO_PRUNE)) atomicMin(&s->ub, (int)best);
    if (sat) flag_saturation(s);
    if (reached) atomicAdd(jcs[cc.jobs[lane]].stx + (blockIdx.x & (NSTRIPE - 1)) * NSTX + ST_STATES, reached);
  }
}

// Zero the label masks of the list just consumed (every item of a vertex read them).
__global__ __launch_bounds__(BLK) void k_cclear(const CohortConst *__restrict__ ccp, const CohortState *__restrict__ csp,
                                                GraphDev g) {
  const CohortConst &cc = *ccp;
  const int cur = csp->par;
  const long long total = (long long)csp->cnt[cur] * cc.NW;
  for (long long i = (long long)blockIdx.x * BLK + threadIdx.x; i < total; i += (long long)gridDim.x * BLK) {
    const long long it = i / cc.NW;
    const int c = (int)(i - it * cc.NW);
    const int32_t *item = cc.items + ((size_t)cur * cc.icap + it) * 2;
    if (item[1] > 0) continue;  // the vertex's first item clears its row
    cc.cact[((size_t)cur * g.n + item[0]) * cc.NW + c] = 0;
  }
}

// Copy the level block back into the members' level-major arenas. A block takes a tile of
// consecutive rows into LDS (one contiguous read), then writes each member's part of the
// tile as one contiguous run of its level block (vt * C words), so both sides stream.
constexpr int SCAT_WORDS = 4096;  // LDS words of a tile: vt = SCAT_WORDS / S rows
static_assert(SCAT_WORDS >= CMAXCH * 64 && CMAXB <= 64, "a scatter tile holds a full row; dbase holds every member");
__global__ __launch_bounds__(BLK) void k_cscatter(const CohortConst *__restrict__ ccp, const JobConst *__restrict__ jcs,
                                                  GraphDev g) {
  const CohortConst &cc = *ccp;
  const int L = cc.L, C = cc.C, S = cc.S, B = cc.B;
  __shared__ int32_t tile[SCAT_WORDS];
  __shared__ int32_t *dbase[64];  // member q's level-L block
  for (int q = threadIdx.x; q < B; q += BLK) {
    const JobConst &c = jcs[cc.jobs[q]];
    dbase[q] = c.T + c.off[L];
  }
  const int VT = SCAT_WORDS / S;  // >= 4: S <= 16 * 64
  const gi32 *CT = as_global(cc.CT);
  for (long long v0 = (long long)blockIdx.x * VT; v0 < g.n; v0 += (long long)gridDim.x * VT) {
    const int nv = (int)min((long long)VT, (long long)g.n - v0);
    __syncthreads();  // the previous tile has been written out (and dbase is visible)
    for (int i = threadIdx.x; i < nv * S; i += BLK) tile[i] = CT[(size_t)v0 * S + i];
    __syncthreads();
    const int per = nv * C;  // words of one member in this tile
    const UDiv dper((uint32_t)per), dc((uint32_t)C);
    for (int i = threadIdx.x; i < B * per; i += BLK) {
      const int q = (int)dper.div((uint32_t)i), r = i - q * per;  // r = vi * C + j
      const int vi = (int)dc.div((uint32_t)r), j = r - vi * C;
      as_global(dbase[q])[(size_t)v0 * C + r] = tile[vi * S + q * C + j];
    }
  }
}

// Phases of the member jobs, set by the cohort driver (the per-query kernels of that phase
// then run on them). mode 1: after the final minimum, REBUILD where a tree is wanted,
// DONE otherwise.
__global__ void k_cset(JobState *jss, const int32_t *jobs, int nj, int phase, int level, int parity, int mode,
                       const JobConst *jcs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nj) return;
  JobState &s = jss[jobs[i]];
  if (mode == 1) {
    s.phase = (s.ans != ~0ull && !(jcs[jobs[i]].flags & JF_NO_TREE)) ? PH_REBUILD : PH_DONE;
    return;
  }
  s.phase = phase;
  s.level = level;
  s.parity = parity;
}

}  // namespace gst
