/* dpbf.c — independent exact GST oracles in plain C. TEST INFRASTRUCTURE ONLY.
 *
 * Written from the published algorithms, not from the reference code, so that they can
 * pin the PrunedDP++ / D-PrunedDP++ restatements:
 *  - orc_dpbf: DPBF (Ding, Yu, Wang, Qin, Zhang, Lin; ICDE 2007). Best-first search over
 *    states (v, p): T(v,p) = min( min_u T(u,p) + w(u,v),  min_{p1 u p2 = p} T(v,p1)+T(v,p2) ),
 *    seeded with T(v,{i}) = 0 for every member v of group i, no pruning; the first
 *    popped state with p = full is optimal.
 *  - orc_dpbf_hop: the diameter-bounded DP computed layer by layer: subsets in order of
 *    popcount, then heights 0..D in order (a grow raises the height by one, a merge at v
 *    of heights d1, d2 with d1 + d2 <= D has height max(d1, d2)); answer = min over v, d
 *    of T(v, full, d). Same transition system as code/D-PrunedDP++/include/CPUHOP.h:192-227
 *    but evaluated in a fixed order instead of by a priority queue.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

#define INF ORC_INF

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* ---- binary min-heap of (cost, v, p) ---- */
typedef struct { int32_t c, v, p; } hnode;
typedef struct { hnode *a; int64_t n, cap; } heap;

static int hless(const hnode *x, const hnode *y) {
  if (x->c != y->c) return x->c < y->c;
  if (x->v != y->v) return x->v < y->v;
  return x->p < y->p;
}
static void hpush(heap *h, int32_t c, int32_t v, int32_t p) {
  if (h->n == h->cap) {
    h->cap = h->cap ? h->cap * 2 : 1024;
    h->a = (hnode *)realloc(h->a, sizeof(hnode) * h->cap);
  }
  int64_t i = h->n++;
  hnode x = {c, v, p};
  while (i > 0) {
    int64_t par = (i - 1) / 2;
    if (!hless(&x, &h->a[par])) break;
    h->a[i] = h->a[par];
    i = par;
  }
  h->a[i] = x;
}
static hnode hpop(heap *h) {
  hnode top = h->a[0], x = h->a[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    const hnode *best = &x;
    if (l < h->n && hless(&h->a[l], best)) { m = l; best = &h->a[l]; }
    if (r < h->n && hless(&h->a[r], best)) { m = r; }
    if (m == i) break;
    h->a[i] = h->a[m];
    i = m;
  }
  if (h->n) h->a[i] = x;
  return top;
}

static int weight_of(const orc_graph *g, int u, int v) {
  int64_t lo = g->beg[u], hi = g->beg[u + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (g->adj[mid] < v) lo = mid + 1; else hi = mid;
  }
  return (lo < g->beg[u + 1] && g->adj[lo] == v) ? g->w[lo] : INF;
}

static int dedup(const int32_t *gids, int k, int32_t *out) {
  int K = 0;
  for (int i = 0; i < k; i++) {
    int seen = 0;
    for (int j = 0; j < K; j++) seen |= out[j] == gids[i];
    if (!seen) out[K++] = gids[i];
  }
  return K;
}

typedef struct { int32_t a, b, w; } edge3;
static int ecmp(const void *x, const void *y) {
  const edge3 *p = (const edge3 *)x, *q = (const edge3 *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  return 0;
}

int orc_dpbf(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *edges,
             int32_t edge_cap, orc_result *res) {
  double t0 = now_s();
  memset(res, 0, sizeof(*res));
  res->dp_cost = res->tree_cost = INF;
  res->root = res->root_d = -1;
  if (k < 1 || k > 14) { res->status = -2; return -2; }
  int32_t q[16];
  for (int i = 0; i < k; i++)
    if (gids[i] < 0 || gids[i] >= g->n_groups) { res->status = -1; return -1; }
  int K = dedup(gids, k, q);
  const int n = g->n, W = 1 << K, R = W - 1;
  size_t NS = (size_t)n * W;
  int32_t *cost = (int32_t *)malloc(sizeof(int32_t) * NS);
  int32_t *bp = (int32_t *)malloc(sizeof(int32_t) * NS); /* >=0 grow from u; -1 seed; <=-2 merge p1=-(bp+2) */
  uint8_t *done = (uint8_t *)calloc(NS, 1);
  for (size_t s = 0; s < NS; s++) cost[s] = INF;
  heap h = {0, 0, 0};
  for (int i = 0; i < K; i++)
    for (int64_t t = g->goffs[q[i]]; t < g->goffs[q[i] + 1]; t++) {
      size_t s = (size_t)g->gverts[t] * W + (1 << i);
      if (cost[s] != 0) { cost[s] = 0; bp[s] = -1; hpush(&h, 0, g->gverts[t], 1 << i); }
    }
  double tb = now_s();
  int rv = -1;
  int64_t process = 0;
  while (h.n) {
    hnode x = hpop(&h);
    size_t s = (size_t)x.v * W + x.p;
    if (done[s] || x.c != cost[s]) continue;
    done[s] = 1;
    process++;
    if (x.p == R) { rv = x.v; break; }
    for (int64_t e = g->beg[x.v]; e < g->beg[x.v + 1]; e++) {
      int u = g->adj[e];
      size_t su = (size_t)u * W + x.p;
      int nc = x.c + g->w[e];
      if (!done[su] && nc < cost[su]) { cost[su] = nc; bp[su] = x.v; hpush(&h, nc, u, x.p); }
    }
    int comp = R & ~x.p;
    for (int p2 = comp; p2 > 0; p2 = (p2 - 1) & comp) {
      size_t s2 = (size_t)x.v * W + p2;
      if (!done[s2]) continue;
      int pu = x.p | p2;
      size_t su = (size_t)x.v * W + pu;
      int nc = x.c + cost[s2];
      if (!done[su] && nc < cost[su]) { cost[su] = nc; bp[su] = -(x.p + 2); hpush(&h, nc, x.v, pu); }
    }
  }
  res->t_search_s = now_s() - tb;
  res->process = process;
  int64_t cnt = 0;
  for (size_t s = 0; s < NS; s++) cnt += cost[s] != INF;
  res->counts = cnt;
  if (rv >= 0) {
    res->root = rv;
    res->dp_cost = cost[(size_t)rv * W + R];
    /* rebuild */
    int64_t cap = 1024, top = 0, ne = 0, ecap = 1024;
    int32_t *st = (int32_t *)malloc(sizeof(int32_t) * 2 * cap);
    edge3 *el = (edge3 *)malloc(sizeof(edge3) * ecap);
    st[0] = rv; st[1] = R; top = 1;
    while (top) {
      top--;
      int v = st[2 * top], p = st[2 * top + 1];
      int b = bp[(size_t)v * W + p];
      if (top + 2 >= cap) { cap *= 2; st = (int32_t *)realloc(st, sizeof(int32_t) * 2 * cap); }
      if (b >= 0) {
        if (ne == ecap) { ecap *= 2; el = (edge3 *)realloc(el, sizeof(edge3) * ecap); }
        el[ne].a = b < v ? b : v; el[ne].b = b < v ? v : b; el[ne].w = weight_of(g, b, v); ne++;
        st[2 * top] = b; st[2 * top + 1] = p; top++;
      } else if (b <= -2) {
        int p1 = -(b + 2);
        st[2 * top] = v; st[2 * top + 1] = p1; top++;
        st[2 * top] = v; st[2 * top + 1] = p - p1; top++;
      }
    }
    qsort(el, ne, sizeof(edge3), ecmp);
    int64_t m = 0, sum = 0;
    for (int64_t i = 0; i < ne; i++)
      if (m == 0 || el[m - 1].a != el[i].a || el[m - 1].b != el[i].b) el[m++] = el[i];
    for (int64_t i = 0; i < m; i++) sum += el[i].w;
    res->n_edges = (int32_t)m;
    if (edges)
      for (int64_t i = 0; i < m && i < edge_cap; i++) {
        edges[3 * i] = el[i].a; edges[3 * i + 1] = el[i].b; edges[3 * i + 2] = el[i].w;
      }
    res->tree_cost = (int32_t)sum;
    res->status = 0;
    free(st);
    free(el);
  } else {
    res->status = 1;
  }
  free(h.a);
  free(cost);
  free(bp);
  free(done);
  res->t_total_s = now_s() - t0;
  return 0;
}

int orc_dpbf_hop(const orc_graph *g, const int32_t *gids, int32_t k, int32_t D,
                 orc_result *res) {
  double t0 = now_s();
  memset(res, 0, sizeof(*res));
  res->dp_cost = res->tree_cost = INF;
  res->root = res->root_d = -1;
  if (k < 1 || k > 14 || D < 0) { res->status = -2; return -2; }
  int32_t q[16];
  for (int i = 0; i < k; i++)
    if (gids[i] < 0 || gids[i] >= g->n_groups) { res->status = -1; return -1; }
  int K = dedup(gids, k, q);
  const int n = g->n, W = 1 << K, R = W - 1, H = D + 1;
  size_t NS = (size_t)n * W * H;
  int32_t *T = (int32_t *)malloc(sizeof(int32_t) * NS);
  for (size_t s = 0; s < NS; s++) T[s] = INF;
#define AT(v, p, d) T[((size_t)(v) * W + (p)) * H + (d)]
  for (int L = 1; L <= K; L++) {
    for (int p = 1; p <= R; p++) {
      if (__builtin_popcount(p) != L) continue;
      if (L == 1) {
        int i = __builtin_ctz(p);
        for (int64_t t = g->goffs[q[i]]; t < g->goffs[q[i] + 1]; t++) AT(g->gverts[t], p, 0) = 0;
      } else {
        for (int v = 0; v < n; v++)
          for (int p1 = (p - 1) & p; p1 > 0; p1 = (p1 - 1) & p) {
            int p2 = p ^ p1;
            for (int d1 = 0; d1 <= D; d1++) {
              int c1 = AT(v, p1, d1);
              if (c1 >= INF) continue;
              for (int d2 = 0; d1 + d2 <= D; d2++) {
                int c2 = AT(v, p2, d2);
                if (c2 >= INF) continue;
                int nd = d1 > d2 ? d1 : d2;
                if (c1 + c2 < AT(v, p, nd)) AT(v, p, nd) = c1 + c2;
              }
            }
          }
      }
      if (L == K) continue; /* growing a full-set tree never lowers the answer */
      for (int d = 0; d < D; d++)
        for (int v = 0; v < n; v++) {
          int c = AT(v, p, d);
          if (c >= INF) continue;
          for (int64_t e = g->beg[v]; e < g->beg[v + 1]; e++) {
            int u = g->adj[e];
            if (c + g->w[e] < AT(u, p, d + 1)) AT(u, p, d + 1) = c + g->w[e];
          }
        }
    }
  }
  int best = INF, bv = -1, bd = -1;
  int64_t cnt = 0;
  for (int v = 0; v < n; v++)
    for (int d = 0; d <= D; d++)
      if (AT(v, R, d) < best) { best = AT(v, R, d); bv = v; bd = d; }
  for (size_t s = 0; s < NS; s++) cnt += T[s] < INF;
#undef AT
  res->counts = cnt;
  res->dp_cost = best;
  res->tree_cost = best;
  res->root = bv;
  res->root_d = bd;
  res->status = bv >= 0 ? 0 : 1;
  free(T);
  res->t_search_s = res->t_total_s = now_s() - t0;
  return 0;
}
