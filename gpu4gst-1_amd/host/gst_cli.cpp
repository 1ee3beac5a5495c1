// gst_cli.cpp — reference-compatible experiment driver over the C ABI.
//
//   gst_cli <type> <path> <name> <T> [D] <start> <end>
//
// Same positional arguments as the reference programs (sh/example.sh:7-42): <path> is
// the data directory (with a trailing '/'), <name> the dataset, <T> the query size k,
// optional <D> the diameter bound, and queries <start>..<end> (inclusive) of
// <path><name><T>.csv are solved. <type>: 2 = one query at a time (TrimCDP /
// D-TrimCDP: exp_GPU_nonHop.h:40-142, exp_GPU_Hop.h:40-122), 3 = all queries through the
// batched persistent path (the role of TrimCDP-WB / D-TrimCDP-WB: GSTnonHop.cu:193-565,
// GPUHop.cu:96-272).
//
// Output: <path>result/exp_GPU{1,2}_{nonHop,Hop}_<name>_T<T>[_D<D>]_<start>-<end>.csv with
// the reference columns task_ID,task,<time>,<cost>,<memory>,counts,process_num. Cost
// uses the reference's no-solution sentinel (0 for TrimCDP, 100000 otherwise); time is
// the DP-loop time in seconds; memory is the reference's state-count proxy
// (finite states + V*2^T*(D+1)).
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gst.h"
#include "../../include/gst_data.h"

static bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

int main(int argc, char **argv) {
  if (argc != 7 && argc != 8) {
    std::fprintf(stderr, "usage: %s <type 2|3> <path> <name> <T> [D] <start> <end>\n", argv[0]);
    return 2;
  }
  const int type = std::atoi(argv[1]);
  const std::string path = argv[2], name = argv[3];
  const int T = std::atoi(argv[4]);
  const bool hop = argc == 8;
  const int D = hop ? std::atoi(argv[5]) : -1;
  const int start = std::atoi(argv[hop ? 6 : 5]), end = std::atoi(argv[hop ? 7 : 6]);
  if (type != 2 && type != 3) { std::fprintf(stderr, "type must be 2 (plain) or 3 (batched)\n"); return 2; }

  gstd_graph *g = nullptr;
  if (file_exists(path + name + "_beg_pos.bin")) g = gstd_read_bin((path + name).c_str());
  else g = gstd_read_in((path + name + ".in").c_str());
  if (!g) { std::fprintf(stderr, "graph: %s\n", gstd_last_error()); return 1; }
  int64_t *goffs = nullptr;
  int32_t *gverts = nullptr;
  const int32_t ng = gstd_read_groups((path + name + ".g").c_str(), &goffs, &gverts);
  if (ng < 0) { std::fprintf(stderr, "groups: %s\n", gstd_last_error()); return 1; }
  int32_t k = 0;
  int32_t *gids = nullptr;
  const int32_t nq = gstd_read_queries((path + name + std::to_string(T) + ".csv").c_str(), &k, &gids);
  if (nq < 0) { std::fprintf(stderr, "queries: %s\n", gstd_last_error()); return 1; }
  std::printf("V %lld groups %d queries %d k %d\n", (long long)gstd_num_vertices(g), ng, nq, k);

  gst_csr csr{(int32_t)gstd_num_vertices(g), gstd_num_arcs(g), gstd_beg_pos(g), gstd_adj(g), gstd_weight(g)};
  gst_groups grp{ng, goffs, gverts};
  gst_opts opts;
  std::memset(&opts, 0, sizeof(opts));
  opts.flags = GST_F_STATS;
  gst_ctx *ctx = nullptr;
  int rc = gst_open(&csr, &grp, &opts, &ctx);
  if (rc) { std::fprintf(stderr, "gst_open: %s: %s\n", gst_strerror(rc), gst_last_error()); return 1; }

  const int s = std::max(0, start), e = std::min(end, nq - 1);
  const std::string tag = type == 2 ? "GPU1" : "GPU2";
  const std::string kind = hop ? "Hop" : "nonHop";
  std::string out = path + "result/exp_" + tag + "_" + kind + "_" + name + "_T" + std::to_string(T) +
                    (hop ? "_D" + std::to_string(D) : std::string()) + "_" + std::to_string(start) + "-" +
                    std::to_string(end) + ".csv";
  mkdir((path + "result").c_str(), 0755);
  FILE *f = std::fopen(out.c_str(), "w");
  if (!f) { std::fprintf(stderr, "cannot write %s\n", out.c_str()); return 1; }
  if (type == 2 && !hop)
    std::fprintf(f, "task_ID,task,GPU1_nonHop_time,GPU_nonHop_cost,GPU_nonHop_memory,counts,process_num\n");
  else
    std::fprintf(f, "task_ID,task,%s_%s_time,%s_%s_cost,%s_%s_memory,counts,process_num\n", tag.c_str(),
                 kind.c_str(), tag.c_str(), kind.c_str(), tag.c_str(), kind.c_str());
  const int none = (type == 2 && !hop) ? 0 : 100000;  // DPQ.cu:571-574; D-TrimCDP DPQ.cu:229-242

  std::vector<gst_result> res(e >= s ? e - s + 1 : 0);
  if (type == 2) {
    for (int i = s; i <= e; i++) {
      rc = gst_solve(ctx, gids + (size_t)i * k, k, D, GST_F_STATS, &res[i - s]);
      if (rc < 0) std::fprintf(stderr, "query %d: %s: %s\n", i, gst_strerror(rc), gst_last_error());
    }
  } else {
    std::vector<gst_query> qs;
    for (int i = s; i <= e; i++) qs.push_back({gids + (size_t)i * k, k, D, GST_F_STATS});
    rc = gst_solve_batch(ctx, qs.data(), (int32_t)qs.size(), res.data());
    if (rc < 0) std::fprintf(stderr, "batch: %s: %s\n", gst_strerror(rc), gst_last_error());
  }
  int failed = 0;
  for (int i = s; i <= e; i++) {
    gst_result &r = res[i - s];
    std::string task;
    for (int j = 0; j < k; j++) task += std::to_string(gids[(size_t)i * k + j]) + " ";
    // errors are not "no solution": their rows carry cost -1 and the exit status is 1
    if (r.status < 0) {
      failed++;
      std::fprintf(stderr, "query %d: %s\n", i, gst_strerror(r.status));
    }
    const int cost = (r.status == GST_OK || r.status == GST_TREE_OMITTED) ? r.cost : r.status < 0 ? -1 : none;
    const long long mem = (long long)r.states_finite + (long long)gstd_num_vertices(g) * (1ll << r.k) * (hop ? D + 1 : 1);
    std::fprintf(f, "%d,%s,%.8f,%d,%lld,%lld,%lld\n", i, task.c_str(), r.t_dp_s, cost, mem,
                 (long long)r.states_finite, (long long)r.states_processed);
    std::printf("%s iteration %d cost %d\n", name.c_str(), i, cost);
    gst_result_free(&r);
  }
  std::fprintf(f, "\n");
  std::fclose(f);
  gst_close(ctx);
  gstd_free(g);
  gstd_free_buf(goffs);
  gstd_free_buf(gverts);
  gstd_free_buf(gids);
  return failed ? 1 : 0;
}
