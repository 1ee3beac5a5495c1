// prunedpp.cpp — CPU restatement of PrunedDP++ and D-PrunedDP++. TEST INFRASTRUCTURE
// ONLY: the parity checker and the reported CPU baseline, never part of the product.
//
// PrunedDP++ follows code/PrunedDP++/include/CPUNONHOP.h:
//   group SSSP via a 0-weight dummy vertex ......... :21-67 (Dijkstra, shortest_paths.h:24-81)
//   virtual group distances ......................... :175-210
//   AllPaths tour table W (with the subset rule) .... :269-386
//   W2 = min_j W(i,j,X) ............................. :568-587
//   LB procedure (tour, tour-2, one-label) .......... :436-496
//   seeding: every non-empty submask of v's groups .. :594-624
//   best-first loop, early return on the full set ... :635-678
//   feasible-solution update + MST post-processing .. :684-726
//   complement merge ("lines 17-19") ................ :733-792
//   grow under cost < best/2 + 2 .................... :795-869
//   merge with popped disjoint sets, 0.667 rule ..... :874-958
//   tree build from type 0/1/2 back-pointers ........ :390-432
//   reported cost = sum of distinct tree edges ...... graph_hash_of_mixed_weighted_sum_of_nw_ec.h:58-87
//   feasibility check, 100000 on failure ............ exp_CPU_nonHop.h:17-49,117-126
// Boost's Fibonacci heap (decrease-key) is replaced by a binary heap with lazy deletion
// plus a per-state "queued priority" array, which reproduces decrease-key semantics; only
// the order among equal priorities differs (ties are broken by (v,p) here).
//
// D-PrunedDP++ follows code/D-PrunedDP++/include/CPUHOP.h:107-288: singleton seeds at
// height 0 (:158-170), lazy-deletion min-heap (:174-185), stop at the first popped full
// state (:186-190), grow to height d+1 while d < D (:192-206), merge with any disjoint p2
// and every d2 <= D-d1 into height max(d1,d2) (:208-227), rebuild (:233-259), reported
// cost = distinct-edge sum with the feasibility check (exp_CPU_Hop.h:142-152).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <numeric>
#include <queue>
#include <thread>
#include <tuple>
#include <unordered_set>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;
const int INF = ORC_INF;

// de-duplicated query (first occurrence kept)
std::vector<int> dedup_query(const int32_t *gids, int k) {
  std::vector<int> q;
  for (int i = 0; i < k; i++)
    if (std::find(q.begin(), q.end(), gids[i]) == q.end()) q.push_back(gids[i]);
  return q;
}

int edge_weight(const orc_graph *g, int u, int v) {
  const int32_t *b = g->adj + g->beg[u], *e = g->adj + g->beg[u + 1];
  const int32_t *it = std::lower_bound(b, e, v);
  if (it == e || *it != v) return INF;
  return g->w[it - g->adj];
}

// Multi-source Dijkstra from the members of one group (the reference adds a dummy vertex
// joined to every member by 0-weight edges; distances are the same). pred[source] = source.
void group_dijkstra(const orc_graph *g, int grp, std::vector<int> &dist, std::vector<int> &pred) {
  const int n = g->n;
  dist.assign(n, INF);
  pred.resize(n);
  std::iota(pred.begin(), pred.end(), 0);
  std::vector<char> done(n, 0);
  typedef std::pair<int, int> E;
  std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
  for (int64_t i = g->goffs[grp]; i < g->goffs[grp + 1]; i++) {
    int s = g->gverts[i];
    if (dist[s] != 0) { dist[s] = 0; pq.push({0, s}); }
  }
  while (!pq.empty()) {
    auto [d, x] = pq.top();
    pq.pop();
    if (done[x] || d != dist[x]) continue;
    done[x] = 1;
    for (int64_t e = g->beg[x]; e < g->beg[x + 1]; e++) {
      int y = g->adj[e], nd = d + g->w[e];
      if (!done[y] && nd < dist[y]) { dist[y] = nd; pred[y] = x; pq.push({nd, y}); }
    }
  }
}

bool group_contains(const orc_graph *g, int grp, int v) {
  const int32_t *b = g->gverts + g->goffs[grp], *e = g->gverts + g->goffs[grp + 1];
  return std::binary_search(b, e, v);
}

// ---- tree helpers: edge sets with distinct (min,max) keys ----
struct EdgeSet {
  std::vector<std::tuple<int, int, int>> e;  // (a<b, w)
  std::vector<int> verts;
  void add_vertex(int v) { verts.push_back(v); }
  void add_edge(int a, int b, int w) {
    if (a > b) std::swap(a, b);
    e.emplace_back(a, b, w);
    verts.push_back(a); verts.push_back(b);
  }
  void normalize() {
    std::sort(e.begin(), e.end());
    // the reference's hash graph keeps one copy per (a,b); a later add overwrites the weight
    std::vector<std::tuple<int, int, int>> out;
    for (auto &t : e) {
      if (!out.empty() && std::get<0>(out.back()) == std::get<0>(t) && std::get<1>(out.back()) == std::get<1>(t))
        out.back() = t;
      else out.push_back(t);
    }
    e.swap(out);
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
  }
  long long sum() const { long long s = 0; for (auto &t : e) s += std::get<2>(t); return s; }
};

// minimum spanning forest weight of an edge set (Kruskal); the MST weight is unique, so
// the reference's Prim start vertex does not change it
long long mst_weight(EdgeSet es) {
  es.normalize();
  std::vector<std::tuple<int, int, int>> byw;
  for (auto &t : es.e) byw.emplace_back(std::get<2>(t), std::get<0>(t), std::get<1>(t));
  std::sort(byw.begin(), byw.end());
  std::vector<int> &vs = es.verts;
  std::vector<int> par(vs.size());
  std::iota(par.begin(), par.end(), 0);
  auto idx = [&](int v) { return (int)(std::lower_bound(vs.begin(), vs.end(), v) - vs.begin()); };
  auto find = [&](int x) { while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; } return x; };
  long long s = 0;
  for (auto &t : byw) {
    int a = find(idx(std::get<1>(t))), b = find(idx(std::get<2>(t)));
    if (a != b) { par[a] = b; s += std::get<0>(t); }
  }
  return s;
}

// feasibility as in exp_CPU_nonHop.h:17-49: one connected component, every group hit
bool feasible(const orc_graph *g, EdgeSet es, const std::vector<int> &q) {
  es.normalize();
  if (es.verts.empty()) return false;
  std::vector<int> &vs = es.verts;
  std::vector<int> par(vs.size());
  std::iota(par.begin(), par.end(), 0);
  auto idx = [&](int v) { return (int)(std::lower_bound(vs.begin(), vs.end(), v) - vs.begin()); };
  auto find = [&](int x) { while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; } return x; };
  int comps = (int)vs.size();
  for (auto &t : es.e) {
    int a = find(idx(std::get<0>(t))), b = find(idx(std::get<1>(t)));
    if (a != b) { par[a] = b; comps--; }
  }
  if (comps != 1) return false;
  for (int grp : q) {
    bool hit = false;
    for (int v : vs) if (group_contains(g, grp, v)) { hit = true; break; }
    if (!hit) return false;
  }
  return true;
}

void emit(const EdgeSet &es_in, int32_t *edges, int32_t cap, orc_result *res) {
  EdgeSet es = es_in;
  es.normalize();
  res->n_edges = (int32_t)es.e.size();
  if (edges)
    for (int i = 0; i < (int)es.e.size() && i < cap; i++) {
      edges[3 * i] = std::get<0>(es.e[i]);
      edges[3 * i + 1] = std::get<1>(es.e[i]);
      edges[3 * i + 2] = std::get<2>(es.e[i]);
    }
}

// ======================================================================= PrunedDP++
struct Prunedpp {
  const orc_graph *g;
  std::vector<int> q;  // group ids, bit i <-> q[i]
  int K, R, W;         // K groups, R = 2^K - 1 (full set), W = 2^K
  int n;
  std::vector<std::vector<int>> dist, pred;  // per bit
  std::vector<std::vector<int>> nonover;     // CPUNONHOP.h:115-141 (j in [1, R-1])
  std::vector<std::vector<int>> uncovered;   // CPUNONHOP.h:145-171 (bit indices)
  std::vector<int> Wt;                       // W[i][j][X] flat, INT_MIN = unset
  std::vector<int> W2;                       // W2[i][X]
  // trees: cost (INT_MAX = absent), back-pointer (type 0: -1; type 1: u>=0; type 2: -(p1+2))
  std::vector<int> cost, bp;
  std::vector<int> qprio;     // queued priority, INT_MAX = not in Q_T
  std::vector<char> popped;   // the reference's set D
  int64_t live = 0, max_live = 0;

  inline int &W3(int i, int j, int X) { return Wt[((size_t)i * K + j) * W + X]; }
  inline int getW(int i, int j, int X) {  // std::unordered_map operator[]: missing -> 0
    int &x = W3(i, j, X);
    if (x == INT_MIN) x = 0;
    return x;
  }

  void all_paths(const std::vector<std::vector<int>> &vd) {  // CPUNONHOP.h:269-386
    Wt.assign((size_t)K * K * W, INT_MIN);
    std::vector<int> qp((size_t)K * K * W, INT_MAX);
    std::vector<char> dn((size_t)K * K * W, 0);
    typedef std::tuple<int, int, int, int> E;  // prio, vi, vj, X
    std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
    auto id = [&](int i, int j, int X) { return ((size_t)i * K + j) * W + X; };
    for (int p = 0; p < K; p++) { pq.push({0, p, p, 0}); qp[id(p, p, 0)] = 0; }
    while (!pq.empty()) {
      auto [c, vi, vj, X] = pq.top();
      pq.pop();
      size_t h = id(vi, vj, X);
      if (dn[h] || qp[h] != c) continue;
      dn[h] = 1;
      qp[h] = INT_MAX;
      W3(vi, vj, X) = c;
      for (int i = 0; i < X; i++)
        if ((i | X) == X) {
          int &x = W3(vi, vj, i);
          if (x == INT_MIN || x > c) x = c;
        }
      for (int p = 0; p < K; p++) {
        int ps = 1 << p;
        if ((ps | X) != X) {
          int Xw = X + ps, cw = c + vd[vj][p];
          size_t h2 = id(vi, p, Xw);
          if (dn[h2]) continue;
          if (cw < qp[h2]) { qp[h2] = cw; pq.push({cw, vi, p, Xw}); }
        }
      }
    }
    W2.assign((size_t)K * W, 0);
    for (int i = 0; i < K; i++)
      for (int X = 1; X <= R; X++) {
        int d = INF;
        for (int j = 0; j < K; j++) d = std::min(d, getW(i, j, X));
        W2[(size_t)i * W + X] = d;
      }
  }

  int lb_procedure(int v, int X, int c) {  // CPUNONHOP.h:436-496
    int lb = 0;
    if (R != X) {
      int Xs = R - X;
      int lb1 = INF, lb2 = -1, lb_one = -1;
      const std::vector<int> &unc = uncovered[X];
      for (int i : unc) {
        int dvi = dist[i][v];
        int xxx = INF;
        for (int j : unc) {
          int dvj = dist[j][v];
          if (xxx > dvj) xxx = dvj;
          int l1 = (dvi + getW(i, j, Xs) + dvj) / 2;
          if (lb1 > l1) lb1 = l1;
        }
        int l2 = (dvi + W2[(size_t)i * W + Xs] + xxx) / 2;
        if (lb2 < l2) lb2 = l2;
        if (lb_one < dvi) lb_one = dvi;
      }
      lb = lb1 < lb2 ? lb2 : lb1;
      if (lb < lb_one) lb = lb_one;
    }
    return c + lb;
  }

  inline size_t S(int v, int p) const { return (size_t)v * W + p; }
  inline int tcost(int v, int p) const { int c = cost[S(v, p)]; return c == INT_MAX ? INF : c; }

  EdgeSet build_tree(int rv, int rp) {  // CPUNONHOP.h:390-432
    EdgeSet es;
    std::vector<std::pair<int, int>> stack{{rv, rp}};
    size_t guard = 0;
    while (!stack.empty()) {
      auto [v, p] = stack.back();
      stack.pop_back();
      es.add_vertex(v);
      int b = bp[S(v, p)];
      if (b == -1) {
      } else if (b >= 0) {
        stack.push_back({b, p});
        es.add_edge(b, v, edge_weight(g, b, v));
      } else {
        int p1 = -(b + 2);
        stack.push_back({v, p1});
        stack.push_back({v, p - p1});
      }
      if (++guard > (size_t)64 * n + 1024) break;  // cannot happen for popped states
    }
    return es;
  }

  struct QE {
    int prio, v, p;
    bool operator>(const QE &o) const {
      return std::tie(prio, v, p) > std::tie(o.prio, o.v, o.p);
    }
  };
  std::priority_queue<QE, std::vector<QE>, std::greater<QE>> Q;
  void push_or_decrease(int v, int p, int prio, int cmp_val) {
    size_t s = S(v, p);
    if (qprio[s] == INT_MAX) {
      qprio[s] = prio; Q.push({prio, v, p});
      if (++live > max_live) max_live = live;
    } else if (cmp_val < qprio[s]) {
      qprio[s] = prio; Q.push({prio, v, p});
    }
  }

  int run(int32_t *edges, int32_t cap, orc_result *res) {
    auto t0 = clk::now();
    n = g->n;
    K = (int)q.size();
    if (K >= 15) return -2;  // CPUNONHOP.h:505-509
    R = (1 << K) - 1;
    W = 1 << K;
    const int safeguard = 2;
    dist.resize(K); pred.resize(K);
    for (int i = 0; i < K; i++) group_dijkstra(g, q[i], dist[i], pred[i]);
    nonover.assign(R + 1, {});
    for (int i = 1; i <= R; i++)
      for (int j = 1; j < R; j++)
        if ((i & j) == 0) nonover[i].push_back(j);
    uncovered.assign(R + 1, {});
    for (int p = 1; p <= R; p++)
      for (int i = 0; i < K; i++)
        if (((1 << i) | p) != p) uncovered[p].push_back(i);
    std::vector<std::vector<int>> vd(K, std::vector<int>(K, INF));
    for (int i = 0; i < K; i++)
      for (int j = i; j < K; j++) {
        int d = INF;
        for (int64_t t = g->goffs[q[j]]; t < g->goffs[q[j] + 1]; t++) d = std::min(d, dist[i][g->gverts[t]]);
        vd[i][j] = vd[j][i] = d;
      }
    all_paths(vd);

    cost.assign((size_t)n * W, INT_MAX);
    bp.assign((size_t)n * W, -1);
    qprio.assign((size_t)n * W, INT_MAX);
    popped.assign((size_t)n * W, 0);
    // seeds: CPUNONHOP.h:594-624
    std::vector<int> vmask(n, 0);
    for (int i = 0; i < K; i++)
      for (int64_t t = g->goffs[q[i]]; t < g->goffs[q[i] + 1]; t++) vmask[g->gverts[t]] |= 1 << i;
    for (int v = 0; v < n; v++) {
      int m = vmask[v];
      for (int p = 1; p <= m; p++)
        if ((p | m) == m) {
          cost[S(v, p)] = 0; bp[S(v, p)] = -1;
          push_or_decrease(v, p, 0, 0);
        }
    }
    int best_cost = INF;
    EdgeSet best_solu;
    int64_t process = 0;
    auto tb = clk::now();
    while (!Q.empty()) {
      QE top = Q.top();
      Q.pop();
      int v = top.v, X = top.p;
      size_t sx = S(v, X);
      if (qprio[sx] != top.prio) continue;  // stale heap entry (decrease-key emulation)
      qprio[sx] = INT_MAX;
      live--;
      process++;
      int vc = cost[sx];
      if (X == R) {  // :658-678
        res->t_search_s = std::chrono::duration<double>(clk::now() - tb).count();
        int64_t cnt = 0;
        for (size_t s = 0; s < cost.size(); s++) cnt += cost[s] != INT_MAX;
        res->counts = cnt;
        res->process = process;
        EdgeSet t = build_tree(v, X);
        res->root = v; res->root_d = -1;
        res->dp_cost = vc;
        bool ok = feasible(g, t, q);
        EdgeSet tn = t; tn.normalize();
        res->tree_cost = ok ? (int)tn.sum() : INF;
        res->status = ok ? 0 : 1;
        emit(t, edges, cap, res);
        res->t_total_s = std::chrono::duration<double>(clk::now() - t0).count();
        return 0;
      }
      popped[sx] = 1;
      // :684-726 feasible solution from T(v,X) plus shortest paths to uncovered groups
      long long fsc = vc;
      for (int i : uncovered[X]) {
        if (dist[i][v] == INF) { fsc = INF; break; }
        fsc += dist[i][v];
      }
      if (fsc < best_cost) {
        EdgeSet fs = build_tree(v, X);
        for (int i : uncovered[X]) {
          int x = v, pr = pred[i][x];
          while (x != pr) {
            fs.add_edge(x, pr, edge_weight(g, x, pr));
            x = pr; pr = pred[i][x];
          }
        }
        best_solu = fs;
        best_cost = (int)mst_weight(fs);
      }
      // :733-792 complement merge
      int Xs = R - X;
      if (popped[S(v, Xs)]) {
        int merged = vc + cost[S(v, Xs)];
        int cur = tcost(v, R);
        if (merged < cur) {
          cost[S(v, R)] = merged; bp[S(v, R)] = -(X + 2);
          int lb = merged;
          if (merged > best_cost + safeguard) continue;
          if (merged < best_cost) { best_cost = merged; best_solu = build_tree(v, R); }
          push_or_decrease(v, R, merged, lb);
        }
        continue;
      }
      // :795-869 grow
      if (vc < best_cost / 2 + safeguard) {
        for (int64_t e = g->beg[v]; e < g->beg[v + 1]; e++) {
          int u = g->adj[e], grow = vc + g->w[e];
          size_t su = S(u, X);
          if (popped[su]) continue;
          if (grow < tcost(u, X)) {
            cost[su] = grow; bp[su] = v;
            int lb = lb_procedure(u, X, grow);
            if (lb > best_cost + safeguard) continue;
            push_or_decrease(u, X, grow, grow);
          }
        }
        // :874-958 merge with popped disjoint sets
        for (int p2 : nonover[X]) {
          if (!popped[S(v, p2)]) continue;
          int pu = X + p2;
          size_t spu = S(v, pu);
          if (popped[spu]) continue;
          int merged = vc + cost[S(v, p2)];
          if (merged < tcost(v, pu)) {
            cost[spu] = merged; bp[spu] = -(X + 2);
            if ((double)merged <= best_cost * 0.667 + safeguard) {
              int lb = lb_procedure(v, pu, merged);
              if (lb > best_cost + safeguard) continue;
              if (pu == R && merged < best_cost) { best_cost = merged; best_solu = build_tree(v, R); }
              push_or_decrease(v, pu, merged, merged);
            }
          }
        }
      }
    }
    // :962-966 no full state popped: return best_solu
    res->t_search_s = std::chrono::duration<double>(clk::now() - tb).count();
    res->process = process;
    int64_t cnt = 0;
    for (size_t s = 0; s < cost.size(); s++) cnt += cost[s] != INT_MAX;
    res->counts = cnt;
    res->root = -1; res->root_d = -1;
    bool ok = feasible(g, best_solu, q);
    EdgeSet tn = best_solu; tn.normalize();
    res->dp_cost = best_cost;
    res->tree_cost = ok ? (int)tn.sum() : INF;
    res->status = ok ? 0 : 1;
    emit(best_solu, edges, cap, res);
    res->t_total_s = std::chrono::duration<double>(clk::now() - t0).count();
    return 0;
  }
};

// ===================================================================== D-PrunedDP++
struct DPrunedpp {
  const orc_graph *g;
  std::vector<int> q;
  int D;
  int run(int32_t *edges, int32_t cap, orc_result *res) {
    auto t0 = clk::now();
    const int n = g->n, K = (int)q.size();
    if (K >= 15 || K < 1) return -2;
    if (D < 0) return -3;
    const int R = (1 << K) - 1, W = 1 << K, H = D + 1;
    auto S = [&](int v, int p, int d) { return ((size_t)v * W + p) * H + d; };
    std::vector<int> cost((size_t)n * W * H, INF);
    // back-pointer: type in {0,1,2}; grow: u; merge: p1, d1, d2 (p2 = p - p1)
    std::vector<int8_t> type((size_t)n * W * H, 0);
    std::vector<int> bpa((size_t)n * W * H, 0);
    std::vector<int16_t> bpd((size_t)n * W * H, 0);  // d1 | d2 << 8
    std::vector<std::vector<int>> nonover(R + 2);      // CPUHOP.h:84-105 (j in [1, R])
    for (int i = 1; i <= R; i++)
      for (int j = 1; j <= R; j++)
        if ((i & j) == 0) nonover[i].push_back(j);
    struct QE {
      int v, p, d, c;
      bool operator>(const QE &o) const { return std::tie(c, v, p, d) > std::tie(o.c, o.v, o.p, o.d); }
    };
    std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
    for (int i = 0; i < K; i++)
      for (int64_t t = g->goffs[q[i]]; t < g->goffs[q[i] + 1]; t++) {
        int v = g->gverts[t];
        size_t s = S(v, 1 << i, 0);
        if (cost[s] != 0) { cost[s] = 0; type[s] = 0; pq.push({v, 1 << i, 0, 0}); }
      }
    int64_t process = 0;
    int rv = -1, rp = 0, rd = 0;
    auto tb = clk::now();
    while (!pq.empty()) {
      process++;
      QE t = pq.top();
      pq.pop();
      int v = t.v, p = t.p, c = t.c, d = t.d;
      if (c != cost[S(v, p, d)]) continue;
      if (p == R) { rv = v; rp = p; rd = d; break; }
      if (d < D)
        for (int64_t e = g->beg[v]; e < g->beg[v + 1]; e++) {
          int u = g->adj[e], len = g->w[e];
          size_t su = S(u, p, d + 1);
          if (cost[su] > c + len) {
            cost[su] = c + len; type[su] = 1; bpa[su] = v;
            pq.push({u, p, d + 1, c + len});
          }
        }
      int p1 = p, d1 = d;
      for (int p2 : nonover[p1])
        for (int d2 = 0; d2 <= D - d1; d2++) {
          int pu = p1 | p2, nd = std::max(d1, d2);
          int mc = c + cost[S(v, p2, d2)];
          size_t sm = S(v, pu, nd);
          if (cost[sm] > mc) {
            cost[sm] = mc; type[sm] = 2; bpa[sm] = p1; bpd[sm] = (int16_t)(d1 | (d2 << 8));
            pq.push({v, pu, nd, mc});
          }
        }
    }
    res->t_search_s = std::chrono::duration<double>(clk::now() - tb).count();
    res->process = process;
    int64_t cnt = 0;
    for (size_t s = 0; s < cost.size(); s++) cnt += cost[s] != INF;
    res->counts = cnt;
    EdgeSet es;
    if (rv >= 0) {
      std::vector<std::tuple<int, int, int>> st{{rv, rp, rd}};
      while (!st.empty()) {
        auto [v, p, d] = st.back();
        st.pop_back();
        es.add_vertex(v);
        size_t s = S(v, p, d);
        if (type[s] == 1) {
          int u = bpa[s];
          es.add_edge(u, v, edge_weight(g, u, v));
          st.push_back({u, p, d - 1});
        } else if (type[s] == 2) {
          int p1 = bpa[s], d1 = bpd[s] & 0xff, d2 = (bpd[s] >> 8) & 0xff;
          st.push_back({v, p1, d1});
          st.push_back({v, p - p1, d2});
        }
      }
      res->dp_cost = cost[S(rv, rp, rd)];
    } else {
      res->dp_cost = INF;
    }
    res->root = rv; res->root_d = rv >= 0 ? rd : -1;
    bool ok = rv >= 0 && feasible(g, es, q);
    EdgeSet tn = es; tn.normalize();
    res->tree_cost = ok ? (int)tn.sum() : INF;
    res->status = ok ? 0 : 1;
    emit(es, edges, cap, res);
    res->t_total_s = std::chrono::duration<double>(clk::now() - t0).count();
    return 0;
  }
};

}  // namespace

extern "C" int orc_dpbf(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *edges,
                        int32_t edge_cap, orc_result *res);

extern "C" {

int orc_prunedpp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *edges,
                 int32_t edge_cap, orc_result *res) {
  std::memset(res, 0, sizeof(*res));
  res->dp_cost = res->tree_cost = INF;
  if (!g || k < 1) { res->status = -1; return -1; }
  for (int i = 0; i < k; i++)
    if (gids[i] < 0 || gids[i] >= g->n_groups) { res->status = -1; return -1; }
  Prunedpp s;
  s.g = g;
  s.q = dedup_query(gids, k);
  int rc = s.run(edges, edge_cap, res);
  if (rc < 0) res->status = rc;
  return rc;
}

int orc_dprunedpp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t D,
                  int32_t *edges, int32_t edge_cap, orc_result *res) {
  std::memset(res, 0, sizeof(*res));
  res->dp_cost = res->tree_cost = INF;
  if (!g || k < 1) { res->status = -1; return -1; }
  for (int i = 0; i < k; i++)
    if (gids[i] < 0 || gids[i] >= g->n_groups) { res->status = -1; return -1; }
  DPrunedpp s;
  s.g = g;
  s.q = dedup_query(gids, k);
  s.D = D;
  int rc = s.run(edges, edge_cap, res);
  if (rc < 0) res->status = rc;
  return rc;
}

int orc_group_sssp(const orc_graph *g, const int32_t *gids, int32_t k, int32_t *dist) {
  std::vector<int> d, p;
  for (int i = 0; i < k; i++) {
    if (gids[i] < 0 || gids[i] >= g->n_groups) return -1;
    group_dijkstra(g, gids[i], d, p);
    std::copy(d.begin(), d.end(), dist + (size_t)i * g->n);
  }
  return 0;
}

int orc_batch(const orc_graph *g, const int32_t *gids, int32_t n_queries, int32_t k,
              int32_t D, int32_t algo, int32_t n_threads, double time_budget_s,
              orc_result *results) {
  for (int i = 0; i < n_queries; i++) { std::memset(&results[i], 0, sizeof(orc_result)); results[i].status = -99; }
  std::atomic<int> next{0}, solved{0};
  auto t0 = clk::now();
  auto worker = [&]() {
    for (;;) {
      if (time_budget_s > 0 && std::chrono::duration<double>(clk::now() - t0).count() > time_budget_s) break;
      int i = next.fetch_add(1);
      if (i >= n_queries) break;
      const int32_t *qq = gids + (size_t)i * k;
      if (algo == 0) orc_prunedpp(g, qq, k, nullptr, 0, &results[i]);
      else if (algo == 1) orc_dprunedpp(g, qq, k, D, nullptr, 0, &results[i]);
      else orc_dpbf(g, qq, k, nullptr, 0, &results[i]);
      solved++;
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < std::max(1, n_threads); t++) th.emplace_back(worker);
  for (auto &t : th) t.join();
  return solved.load();
}

}  // extern "C"
