// gst_data.cpp — synthetic graph generators and reference file formats (libgstdata.so).
//
// Graph shape conventions (simple undirected, deduplicated, no self loops, largest
// connected component, relabelled in original-id order) and the reader semantics of
// the reference's text/binary formats:
//   .in reader   code/PrunedDP++/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:15-66
//   .g reader    same file :116-171  (line "g<n>:..." stored at index n-1, members sorted)
//   .csv reader  same file :172-218
//   .bin loader  code/D-TrimCDP-WB/include/graph.hpp:7-118 (int64 files, V = size/8 - 1)
#include "../../include/gst_data.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <parallel/algorithm>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

struct gstd_graph {
  int64_t n = 0;
  std::vector<int64_t> beg;  // n+1
  std::vector<int32_t> adj;  // m arcs
  std::vector<int32_t> w;    // m arcs
};

namespace {

thread_local std::string g_err;
void set_err(const std::string &s) { g_err = s; }

// ---- PRNG: splitmix64 seeding + xoshiro256** (Blackman & Vigna, 2018) ----
inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed, uint64_t stream = 0) {
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
    for (auto &v : s) v = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  inline double uniform() { return (next() >> 11) * 0x1.0p-53; }
  // unbiased integer in [0, n) (Lemire's multiply-shift with rejection)
  inline uint64_t below(uint64_t n) {
    unsigned __int128 m = (unsigned __int128)next() * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
      uint64_t t = (0 - n) % n;
      while (l < t) { m = (unsigned __int128)next() * n; l = (uint64_t)m; }
    }
    return (uint64_t)(m >> 64);
  }
};

int n_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(h, 16u);
}

template <class F>
void parallel_for_chunks(int64_t n_chunks, F f) {
  int T = n_threads();
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([&]() {
      for (;;) {
        int64_t c = next.fetch_add(1);
        if (c >= n_chunks) break;
        f(c);
      }
    });
  for (auto &x : th) x.join();
}

inline uint64_t edge_key(uint32_t a, uint32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)a << 32) | b;
}

void sort_unique(std::vector<uint64_t> &keys) {
  __gnu_parallel::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
}

// Restrict a deduplicated undirected edge-key list on n vertices to its largest
// connected component, relabel in original-id order, attach weights, build CSR.
gstd_graph *finish_graph(int64_t n, std::vector<uint64_t> &keys, int wlo, int whi,
                         uint64_t seed) {
  std::vector<int32_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int32_t x) {
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
  };
  for (uint64_t k : keys) {
    int32_t a = find((int32_t)(k >> 32)), b = find((int32_t)(k & 0xffffffffu));
    if (a != b) { if (a < b) parent[b] = a; else parent[a] = b; }
  }
  std::vector<int64_t> csize(n, 0);
  for (int64_t v = 0; v < n; v++) csize[find((int32_t)v)]++;
  int32_t best = 0;
  for (int64_t v = 0; v < n; v++)
    if (csize[v] > csize[best]) best = (int32_t)v;
  std::vector<int32_t> relabel(n, -1);
  int64_t nn = 0;
  for (int64_t v = 0; v < n; v++)
    if (find((int32_t)v) == best) relabel[v] = (int32_t)nn++;
  // keep edges inside the component (keys stay sorted after relabel: monotone map)
  std::vector<uint64_t> kept;
  kept.reserve(keys.size());
  for (uint64_t k : keys) {
    int32_t a = relabel[k >> 32], b = relabel[k & 0xffffffffu];
    if (a >= 0 && b >= 0) kept.push_back(((uint64_t)a << 32) | (uint32_t)b);
  }
  keys.swap(kept);
  std::vector<uint64_t>().swap(kept);
  const int64_t E = (int64_t)keys.size();
  // weights: a stream separate from the topology stream, indexed by sorted edge order
  std::vector<int32_t> ew(E);
  const int64_t CH = 1 << 20;
  const uint64_t span = (uint64_t)(whi - wlo + 1);
  parallel_for_chunks((E + CH - 1) / CH, [&](int64_t c) {
    Rng r(seed, 0x57E1647ull + (uint64_t)c);
    int64_t e1 = std::min(E, (c + 1) * CH);
    for (int64_t e = c * CH; e < e1; e++) ew[e] = wlo + (int32_t)r.below(span);
  });
  auto *g = new gstd_graph();
  g->n = nn;
  g->beg.assign(nn + 1, 0);
  for (uint64_t k : keys) { g->beg[(k >> 32) + 1]++; g->beg[(k & 0xffffffffu) + 1]++; }
  for (int64_t v = 0; v < nn; v++) g->beg[v + 1] += g->beg[v];
  g->adj.resize(2 * E);
  g->w.resize(2 * E);
  std::vector<int64_t> pos(g->beg.begin(), g->beg.end() - 1);
  // Keys are sorted by (a,b) with a<b, so for a vertex x every key (a,x) precedes every
  // key (x,b): appending arcs in key order leaves each adjacency list ascending. The
  // pass below only verifies that (and sorts a list if it ever is not).
  for (int64_t e = 0; e < E; e++) {
    uint32_t a = keys[e] >> 32, b = keys[e] & 0xffffffffu;
    g->adj[pos[a]] = (int32_t)b; g->w[pos[a]++] = ew[e];
    g->adj[pos[b]] = (int32_t)a; g->w[pos[b]++] = ew[e];
  }
  parallel_for_chunks((nn + 4095) / 4096, [&](int64_t c) {
    std::vector<std::pair<int32_t, int32_t>> tmp;
    int64_t v1 = std::min<int64_t>(nn, (c + 1) * 4096);
    for (int64_t v = c * 4096; v < v1; v++) {
      int64_t b0 = g->beg[v], b1 = g->beg[v + 1];
      bool sorted = true;
      for (int64_t i = b0 + 1; i < b1; i++)
        if (g->adj[i - 1] > g->adj[i]) { sorted = false; break; }
      if (sorted) continue;
      tmp.clear();
      for (int64_t i = b0; i < b1; i++) tmp.push_back({g->adj[i], g->w[i]});
      std::sort(tmp.begin(), tmp.end());
      for (int64_t i = b0; i < b1; i++) { g->adj[i] = tmp[i - b0].first; g->w[i] = tmp[i - b0].second; }
    }
  });
  return g;
}

// Walker/Vose alias table over non-negative weights.
struct Alias {
  std::vector<double> prob;
  std::vector<uint32_t> alias;
  explicit Alias(const std::vector<double> &wt) {
    size_t n = wt.size();
    prob.resize(n); alias.resize(n);
    double sum = 0; for (double x : wt) sum += x;
    std::vector<double> p(n);
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; i++) {
      p[i] = wt[i] * n / sum;
      (p[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t s = small.back(); small.pop_back();
      uint32_t l = large.back();
      prob[s] = p[s]; alias[s] = l;
      p[l] = (p[l] + p[s]) - 1.0;
      if (p[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
    for (uint32_t l : large) { prob[l] = 1.0; alias[l] = l; }
    for (uint32_t s : small) { prob[s] = 1.0; alias[s] = s; }
  }
  inline uint32_t draw(Rng &r) const {
    uint64_t x = r.next();
    uint32_t i = (uint32_t)(((x >> 32) * (uint64_t)prob.size()) >> 32);
    double u = (x & 0xffffffffu) * 0x1.0p-32;
    return u < prob[i] ? i : alias[i];
  }
};

}  // namespace

extern "C" {

const char *gstd_last_error(void) { return g_err.c_str(); }

gstd_graph *gstd_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                          int wlo, int whi, uint64_t seed) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || wlo < 1 || whi < wlo) {
    set_err("gstd_gen_rmat: bad arguments");
    return nullptr;
  }
  const int64_t n = 1ll << scale, M = (int64_t)edge_factor << scale;
  std::vector<uint64_t> keys(M);
  const int64_t CH = 1 << 16;
  std::vector<uint8_t> drop(M, 0);
  parallel_for_chunks((M + CH - 1) / CH, [&](int64_t ch) {
    Rng r(seed, (uint64_t)ch);
    int64_t e1 = std::min(M, (ch + 1) * CH);
    for (int64_t e = ch * CH; e < e1; e++) {
      uint32_t u = 0, v = 0;
      for (int l = 0; l < scale; l++) {
        double x = r.uniform();
        int q = x < a ? 0 : x < a + b ? 1 : x < a + b + c ? 2 : 3;
        u = (u << 1) | (q >> 1);
        v = (v << 1) | (q & 1);
      }
      if (u == v) drop[e] = 1;
      keys[e] = edge_key(u, v);
    }
  });
  int64_t j = 0;
  for (int64_t e = 0; e < M; e++)
    if (!drop[e]) keys[j++] = keys[e];
  keys.resize(j);
  sort_unique(keys);
  return finish_graph(n, keys, wlo, whi, seed ^ 0xA5A5A5A5ull);
}

gstd_graph *gstd_gen_chunglu(int64_t n, int64_t m, double beta, int wlo, int whi,
                             uint64_t seed) {
  if (n < 2 || n > (1ll << 31) - 1 || m < 1 || beta <= 1.0 || wlo < 1 || whi < wlo ||
      m > n * (n - 1) / 2) {
    set_err("gstd_gen_chunglu: bad arguments");
    return nullptr;
  }
  std::vector<double> wt(n);
  const double ex = -1.0 / (beta - 1.0);
  for (int64_t i = 0; i < n; i++) wt[i] = std::pow((double)(i + 1), ex);
  Alias al(wt);
  std::vector<double>().swap(wt);
  std::vector<uint64_t> keys;
  keys.reserve((size_t)(m * 1.15) + 1024);
  uint64_t round = 0;
  while ((int64_t)keys.size() < m && round < 64) {
    const int64_t need = m - (int64_t)keys.size();
    const int64_t draw = need + need / 8 + 1024;
    const size_t base = keys.size();
    keys.resize(base + draw);
    const int64_t CH = 1 << 16;
    parallel_for_chunks((draw + CH - 1) / CH, [&](int64_t ch) {
      Rng r(seed, (round << 40) + (uint64_t)ch);
      int64_t e1 = std::min(draw, (ch + 1) * CH);
      for (int64_t e = ch * CH; e < e1; e++) {
        uint32_t u = al.draw(r), v = al.draw(r);
        keys[base + e] = u == v ? ~0ull : edge_key(u, v);
      }
    });
    sort_unique(keys);
    if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
    round++;
  }
  if ((int64_t)keys.size() > m) {
    // drop the surplus uniformly at random (deterministic), keeping sorted order
    int64_t surplus = (int64_t)keys.size() - m;
    Rng r(seed, 0xD409ull);
    std::vector<uint8_t> del(keys.size(), 0);
    while (surplus > 0) {
      uint64_t i = r.below(keys.size());
      if (!del[i]) { del[i] = 1; surplus--; }
    }
    int64_t j = 0;
    for (size_t i = 0; i < keys.size(); i++)
      if (!del[i]) keys[j++] = keys[i];
    keys.resize(j);
  }
  return finish_graph(n, keys, wlo, whi, seed ^ 0x5A5A5A5Aull);
}

gstd_graph *gstd_from_edges(int64_t n, int64_t n_edges, const int32_t *u, const int32_t *v,
                            const int32_t *w) {
  if (n < 0 || n_edges < 0) { set_err("gstd_from_edges: bad sizes"); return nullptr; }
  auto *g = new gstd_graph();
  g->n = n;
  g->beg.assign(n + 1, 0);
  for (int64_t e = 0; e < n_edges; e++) {
    if (u[e] < 0 || u[e] >= n || v[e] < 0 || v[e] >= n) {
      delete g; set_err("gstd_from_edges: vertex id out of range"); return nullptr;
    }
    g->beg[u[e] + 1]++; g->beg[v[e] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) g->beg[i + 1] += g->beg[i];
  g->adj.resize(2 * n_edges); g->w.resize(2 * n_edges);
  std::vector<int64_t> pos(g->beg.begin(), g->beg.end() - 1);
  for (int64_t e = 0; e < n_edges; e++) {
    g->adj[pos[u[e]]] = v[e]; g->w[pos[u[e]]++] = w[e];
    g->adj[pos[v[e]]] = u[e]; g->w[pos[v[e]]++] = w[e];
  }
  std::vector<std::pair<int32_t, int32_t>> tmp;
  for (int64_t x = 0; x < n; x++) {
    tmp.clear();
    for (int64_t i = g->beg[x]; i < g->beg[x + 1]; i++) tmp.push_back({g->adj[i], g->w[i]});
    std::sort(tmp.begin(), tmp.end());
    for (int64_t i = g->beg[x]; i < g->beg[x + 1]; i++) {
      g->adj[i] = tmp[i - g->beg[x]].first; g->w[i] = tmp[i - g->beg[x]].second;
    }
  }
  return g;
}

int64_t gstd_num_vertices(const gstd_graph *g) { return g ? g->n : -1; }
int64_t gstd_num_arcs(const gstd_graph *g) { return g ? (int64_t)g->adj.size() : -1; }
const int64_t *gstd_beg_pos(const gstd_graph *g) { return g->beg.data(); }
const int32_t *gstd_adj(const gstd_graph *g) { return g->adj.data(); }
const int32_t *gstd_weight(const gstd_graph *g) { return g->w.data(); }
void gstd_free(gstd_graph *g) { delete g; }
void gstd_free_buf(void *p) { std::free(p); }

int gstd_write_in(const gstd_graph *g, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) { set_err(std::string("cannot open ") + path); return -1; }
  std::fprintf(f, "%lld %lld\n", (long long)g->n, (long long)(g->adj.size() / 2));
  for (int64_t x = 0; x < g->n; x++)
    for (int64_t i = g->beg[x]; i < g->beg[x + 1]; i++)
      if (g->adj[i] > x) std::fprintf(f, "%lld %d %d\n", (long long)x, g->adj[i], g->w[i]);
  std::fclose(f);
  return 0;
}

static int write_i64(const std::string &p, const int64_t *d, size_t n) {
  FILE *f = std::fopen(p.c_str(), "wb");
  if (!f) { set_err("cannot open " + p); return -1; }
  size_t wr = std::fwrite(d, sizeof(int64_t), n, f);
  std::fclose(f);
  if (wr != n) { set_err("short write " + p); return -1; }
  return 0;
}

int gstd_write_bin(const gstd_graph *g, const char *prefix) {
  std::string pre(prefix);
  std::vector<int64_t> tmp(g->adj.size());
  if (write_i64(pre + "_beg_pos.bin", g->beg.data(), g->beg.size())) return -1;
  for (size_t i = 0; i < tmp.size(); i++) tmp[i] = g->adj[i];
  if (write_i64(pre + "_csr.bin", tmp.data(), tmp.size())) return -1;
  for (size_t i = 0; i < tmp.size(); i++) tmp[i] = g->w[i];
  return write_i64(pre + "_weight.bin", tmp.data(), tmp.size());
}

gstd_graph *gstd_read_in(const char *path) {
  std::ifstream in(path);
  if (!in) { set_err(std::string("Unable to open file ") + path); return nullptr; }
  std::string line;
  if (!std::getline(in, line)) { set_err("empty .in file"); return nullptr; }
  long long V = -1;
  { std::istringstream ss(line); ss >> V; }
  if (V < 0) { set_err("bad .in header"); return nullptr; }
  std::vector<int32_t> u, v, w;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    long long a, b, c;
    if (ss >> a >> b >> c) {
      std::string extra;
      if (ss >> extra) continue;  // the reference keeps only 3-field lines
      u.push_back((int32_t)a); v.push_back((int32_t)b); w.push_back((int32_t)c);
    }
  }
  return gstd_from_edges(V, (int64_t)u.size(), u.data(), v.data(), w.data());
}

static bool read_i64(const std::string &p, std::vector<int64_t> &out) {
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) { set_err("cannot open " + p); return false; }
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(sz / 8);
  size_t rd = std::fread(out.data(), 8, out.size(), f);
  std::fclose(f);
  if (rd != out.size()) { set_err("short read " + p); return false; }
  return true;
}

gstd_graph *gstd_read_bin(const char *prefix) {
  std::string pre(prefix);
  std::vector<int64_t> beg, adj, w;
  if (!read_i64(pre + "_beg_pos.bin", beg) || !read_i64(pre + "_csr.bin", adj) ||
      !read_i64(pre + "_weight.bin", w))
    return nullptr;
  if (beg.empty() || (int64_t)adj.size() != beg.back() || adj.size() != w.size()) {
    set_err("inconsistent .bin sizes");
    return nullptr;
  }
  auto *g = new gstd_graph();
  g->n = (int64_t)beg.size() - 1;
  g->beg = beg;
  g->adj.assign(adj.begin(), adj.end());
  g->w.assign(w.begin(), w.end());
  return g;
}

int gstd_write_groups(const char *path, int32_t n_groups, const int64_t *offs,
                      const int32_t *verts) {
  FILE *f = std::fopen(path, "w");
  if (!f) { set_err(std::string("cannot open ") + path); return -1; }
  for (int32_t g = 0; g < n_groups; g++) {
    std::fprintf(f, "g%d:", g + 1);
    for (int64_t i = offs[g]; i < offs[g + 1]; i++)
      std::fprintf(f, i + 1 < offs[g + 1] ? "%d " : "%d", verts[i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  return 0;
}

int32_t gstd_read_groups(const char *path, int64_t **offs_out, int32_t **verts_out) {
  std::ifstream in(path);
  if (!in) { set_err(std::string("Unable to open file ") + path); return -1; }
  std::vector<std::vector<int32_t>> groups;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    size_t colon = line.find(':');
    if (line[0] != 'g' || colon == std::string::npos) { set_err("malformed .g line: " + line); return -2; }
    long gid;
    try { gid = std::stol(line.substr(1, colon - 1)); } catch (...) { set_err("malformed .g id: " + line); return -2; }
    if (gid < 1) { set_err("group ids are 1-based: " + line); return -2; }
    if ((long)groups.size() < gid) groups.resize(gid);
    std::istringstream ss(line.substr(colon + 1));
    std::string tok;
    std::vector<int32_t> mem;
    while (ss >> tok) {
      try { mem.push_back((int32_t)std::stol(tok)); } catch (...) { set_err("malformed .g member: " + line); return -2; }
    }
    std::sort(mem.begin(), mem.end());
    groups[gid - 1] = mem;
  }
  int64_t total = 0;
  for (auto &gv : groups) total += (int64_t)gv.size();
  auto *offs = (int64_t *)std::malloc(sizeof(int64_t) * (groups.size() + 1));
  auto *verts = (int32_t *)std::malloc(sizeof(int32_t) * (total > 0 ? total : 1));
  offs[0] = 0;
  for (size_t g = 0; g < groups.size(); g++) {
    std::copy(groups[g].begin(), groups[g].end(), verts + offs[g]);
    offs[g + 1] = offs[g] + (int64_t)groups[g].size();
  }
  *offs_out = offs;
  *verts_out = verts;
  return (int32_t)groups.size();
}

int gstd_write_queries(const char *path, int32_t n_queries, int32_t k, const int32_t *gids) {
  FILE *f = std::fopen(path, "w");
  if (!f) { set_err(std::string("cannot open ") + path); return -1; }
  for (int32_t q = 0; q < n_queries; q++) {
    for (int32_t j = 0; j < k; j++) std::fprintf(f, j + 1 < k ? "%d " : "%d", gids[q * k + j]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  return 0;
}

int32_t gstd_read_queries(const char *path, int32_t *k_out, int32_t **gids_out) {
  std::ifstream in(path);
  if (!in) { set_err(std::string("Unable to open file ") + path); return -1; }
  std::vector<int32_t> all;
  std::string line;
  int32_t k = -1, rows = 0;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok;
    std::vector<int32_t> row;
    while (ss >> tok) {
      try { row.push_back((int32_t)std::stol(tok)); } catch (...) { set_err("malformed query line: " + line); return -2; }
    }
    if (row.empty()) continue;
    if (k < 0) k = (int32_t)row.size();
    if ((int32_t)row.size() != k) { set_err("query rows of unequal length"); return -2; }
    all.insert(all.end(), row.begin(), row.end());
    rows++;
  }
  auto *out = (int32_t *)std::malloc(sizeof(int32_t) * (all.empty() ? 1 : all.size()));
  std::copy(all.begin(), all.end(), out);
  *k_out = k < 0 ? 0 : k;
  *gids_out = out;
  return rows;
}

}  // extern "C"
