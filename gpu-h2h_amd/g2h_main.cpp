// G2H -- the reference's command line (main.cu:244-307) over libh2h's C ABI.
//
//   ./G2H GraphName PartitionHeight ChangeToGlobalHeight threadNum
//         partitionMethod useGPUContract parallelFine useGPUConstruct
//
// Reads ./Graph/USA/<G>Edge.txt and ./Partition/<G>/<method>{ID,Cut}.txt
// (method 1 Latitude, 2 Minimum, 3 metis, 4 scotch; main.cu:262-292), runs
// contraction_noHub / constructLabel_noHub / printInfo / Query_noHub
// (main.cu:149-240) and appends the reference's CSV rows:
//   ./ConstructInfo/<G>/contract.csv and constructInfo.csv (31 columns),
//   ./QueryResult/<G>/random.csv (23 columns),
//   ./<Method>PartitionRate.csv,
// and, for GraphName NY with ./Query/NY/{normal,sparse}.txt present, the
// real-query row (answerRealNYQuery_noHub) in
//   ./QueryResult/NY/PTHeight-<H>-ChangeHeight-<C>-real.csv (19 columns).
// Every phase runs on the GPU: useGPUContract picks the elimination order
// (0 = partition_TD_T's, 1 = makeCHPerHeight_D_2's), parallelFine picks the
// build's work split (0/1 per node, 2 per bag entry, 3 batched), and
// useGPUConstruct = 0 is accepted but the build still runs on the GPU.
// H2H_LABEL_KIND=3 reproduces the as-shipped L3 labels (default 4, exact).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../include/h2h.h"

namespace {

bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}

void mkdirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); i++) {
    cur += path[i];
    if (path[i] == '/') mkdir(cur.c_str(), 0755);
  }
  mkdir(path.c_str(), 0755);
}

int check(h2h_ctx* ctx, h2h_status st, const char* what) {
  if (st != H2H_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, st, ctx ? h2h_last_error(ctx) : "");
    return 1;
  }
  return 0;
}

// Up to `limit` (s, t, other) triples of a query file; fewer when it ends
// early (uploadRealQuery, Query.cu:21-44, reads 1e5 from each file)
void read_triples(const std::string& path, int64_t limit, std::vector<int32_t>& x, std::vector<int32_t>& y) {
  std::ifstream f(path);
  long long s, t, other;
  for (int64_t i = 0; i < limit && (f >> s >> t >> other); i++) {
    x.push_back((int32_t)s);
    y.push_back((int32_t)t);
  }
}

// answerRealNYQuery_noHub (Query.cu:626-760): the queries of
// ./Query/NY/normal.txt and ./Query/NY/sparse.txt (vertex ids as written),
// timed on the plain kernels, one row appended to
// ./QueryResult/NY/PTHeight-<H>-ChangeHeight-<C>-real.csv under the
// reference's 19-column header.  Skipped when either file is missing.
int real_queries(h2h_ctx* ctx, const h2h_info& info, int C, int label_kind) {
  const std::string normal = "./Query/NY/normal.txt", sparse = "./Query/NY/sparse.txt";
  if (!file_exists(normal) || !file_exists(sparse)) return 0;
  std::vector<int32_t> x, y;
  read_triples(normal, 100000, x, y);
  read_triples(sparse, 100000, x, y);
  const int64_t nq = (int64_t)x.size();
  for (int64_t i = 0; i < nq; i++)
    if (x[i] < 0 || x[i] >= info.n || y[i] < 0 || y[i] >= info.n) {
      std::fprintf(stderr, "real queries: pair %lld (%d, %d) is outside the graph's %lld vertices\n", (long long)i,
                   x[i], y[i], (long long)info.n);
      return 1;
    }
  std::vector<int32_t> out(std::max<int64_t>(nq, 1));
  double t_nolca = 0, t_lca = 0;
  if (nq > 0) {
    if (check(ctx, h2h_query(ctx, H2H_QUERY_NOLCA, H2H_EXEC_PLAIN, x.data(), y.data(), out.data(), nq, &t_nolca),
              "real queries"))
      return 1;
    if (label_kind == H2H_LABEL_L4 &&
        check(ctx, h2h_query(ctx, H2H_QUERY_LCA, H2H_EXEC_PLAIN, x.data(), y.data(), out.data(), nq, &t_lca),
              "real queries"))
      return 1;
  }
  std::printf("real queries (%lld): noLCA %.0f us, LCA %.0f us\n", (long long)nq, t_nolca, t_lca);
  mkdirs("./QueryResult/NY");
  std::ofstream q("./QueryResult/NY/PTHeight-" + std::to_string(info.partition_height) + "-ChangeHeight-" +
                      std::to_string(C) + "-real.csv",
                  std::ios::app);
  // the reference writes the header before every row, and 10000 as the
  // query number (its loop starts at i = queryMaxSize but prints 10000 + 0 * queryStep)
  q << "queryNumber,sequence,multiThread-4,multiThread-6,multiThread-8,multiThread-10,multiThread-12,"
       "multiThread-14,multiThread-16,GPUnoLCA-64,GPUwithLCA-64,GPUnoLCA-128,GPUwithLCA-128,GPUnoLCA-256,"
       "GPUwithLCA-256,GPUnoLCA-512,GPUwithLCA-512,GPUnoLCA-1024,GPUwithLCA-1024\n";
  q << 10000 << ",-1,-1,-1,-1,-1,-1,-1,-1";
  for (int b = 0; b < 5; b++) q << "," << (long long)t_nolca << "," << (long long)t_lca;
  q << "\n";
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 9) {
    std::fprintf(stderr,
                 "usage: G2H GraphName PartitionHeight ChangeToGlobalHeight threadNum partitionMethod "
                 "useGPUContract parallelFine useGPUConstruct\n");
    return 0;  // the reference returns 0 on a wrong argument count (main.cu:246-247)
  }
  const std::string graph = argv[1];
  const int H = std::atoi(argv[2]);
  const int C = std::atoi(argv[3]);
  const int threads = std::atoi(argv[4]);
  const int method = std::atoi(argv[5]);
  const int contract_device = std::atoi(argv[6]);
  const int parallel_fine = std::atoi(argv[7]);
  const int construct_device = std::atoi(argv[8]);
  const char* kind_env = std::getenv("H2H_LABEL_KIND");
  const int label_kind = (kind_env && std::atoi(kind_env) == 3) ? H2H_LABEL_L3 : H2H_LABEL_L4;

  const std::string edge = "./Graph/USA/" + graph + "Edge.txt";
  std::string mname = "Latitude";
  if (method == 2) mname = "Minimum";
  else if (method == 3) mname = "metis";
  else if (method == 4) mname = "scotch";
  const std::string cut = "./Partition/" + graph + "/" + mname + "Cut.txt";
  const std::string idf = "./Partition/" + graph + "/" + mname + "ID.txt";
  std::printf("%s\n%s\n%s\n", edge.c_str(), cut.c_str(), idf.c_str());

  h2h_ctx* ctx = nullptr;
  if (check(nullptr, h2h_create(&ctx, 0), "h2h_create")) return 1;
  auto t0 = std::chrono::steady_clock::now();
  if (check(ctx, h2h_load_files(ctx, edge.c_str(), idf.c_str(), cut.c_str(), H), "load")) return 1;
  auto us_since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t).count();
  };
  std::printf("graph loaded in %.0f us\n", us_since(t0));

  // contraction_noHub (main.cu:149-158)
  double contract_us = 0;
  if (check(ctx, h2h_contract(ctx, C, contract_device ? H2H_CONTRACT_GPU_ORDER : H2H_CONTRACT_CPU_ORDER,
                              threads, &contract_us),
            "contraction"))
    return 1;
  h2h_info info;
  h2h_get_info(ctx, &info);
  std::printf("contraction: %.0f us on the GPU, tree height %lld, %lld label entries\n", contract_us,
              (long long)info.tree_height, (long long)info.label_entries);
  const std::string cdir = "./ConstructInfo/" + graph;
  mkdirs(cdir);
  h2h_write_construct_info(ctx, (cdir + "/contract.csv").c_str(), graph.c_str(), method, threads, contract_device,
                           parallel_fine, construct_device);
  {
    // <Method>PartitionRate.csv row (GraphPartition.cu:252-265)
    std::vector<int32_t> nan(3 * ((1 << info.partition_height) - 1));
    std::vector<double> frac(info.partition_height);
    h2h_ctx* host = nullptr;
    if (h2h_create(&host, -1) == H2H_OK && h2h_load_files(host, edge.c_str(), idf.c_str(), cut.c_str(), H) == H2H_OK &&
        h2h_partition_rank_tree(host, nan.data(), nullptr, frac.data()) == H2H_OK) {
      std::ofstream rate(mname + "PartitionRate.csv", std::ios::app);
      rate << graph << ",";
      for (int i = (int)frac.size() - 1; i >= 0; i--) rate << frac[i] << ",";
      rate << "\n";
    }
    h2h_destroy(host);
  }

  // constructLabel_noHub (main.cu:160-188)
  if (construct_device == 0) std::printf("note: labels are built on the GPU (no CPU construction in this build)\n");
  const int build_method = parallel_fine == 2 ? H2H_METHOD_LL : (parallel_fine == 3 ? H2H_METHOD_BATCHED : H2H_METHOD_PF);
  double label_us = 0;
  if (check(ctx, h2h_build_labels(ctx, label_kind, build_method, &label_us), "label construction")) return 1;
  std::printf("labels (L%d): %.0f us on the GPU\n", label_kind, label_us);
  h2h_write_construct_info(ctx, (cdir + "/constructInfo.csv").c_str(), graph.c_str(), method, threads,
                           contract_device, parallel_fine, construct_device);

  // Query_noHub / answerRandomQuery_noHub (main.cu:218-240, Query.cu:479-624)
  const int64_t nq = 1000000;  // queryMaxSize (Graph_D.cuh:500)
  std::vector<int32_t> x(nq), y(nq), out(nq), out2(nq);
  std::mt19937_64 rng((uint64_t)std::chrono::steady_clock::now().time_since_epoch().count());
  std::uniform_int_distribution<int32_t> pick(0, (int32_t)info.n - 1);
  for (int64_t i = 0; i < nq; i++) {
    x[i] = pick(rng);
    y[i] = pick(rng);
  }
  double t_nolca = 0, t_lca = 0, t_nolca_b = 0, t_lca_b = 0;
  if (check(ctx, h2h_query(ctx, H2H_QUERY_NOLCA, H2H_EXEC_PLAIN, x.data(), y.data(), out.data(), nq, &t_nolca), "query"))
    return 1;
  if (label_kind == H2H_LABEL_L4 &&
      check(ctx, h2h_query(ctx, H2H_QUERY_LCA, H2H_EXEC_PLAIN, x.data(), y.data(), out2.data(), nq, &t_lca), "query"))
    return 1;
  h2h_query(ctx, H2H_QUERY_NOLCA, H2H_EXEC_BATCHED, x.data(), y.data(), out2.data(), nq, &t_nolca_b);
  if (label_kind == H2H_LABEL_L4) h2h_query(ctx, H2H_QUERY_LCA, H2H_EXEC_BATCHED, x.data(), y.data(), out2.data(), nq, &t_lca_b);
  std::printf("queries (1e6): noLCA plain %.0f us, batched %.0f us; LCA plain %.0f us, batched %.0f us\n", t_nolca,
              t_nolca_b, t_lca, t_lca_b);
  const std::string qdir = "./QueryResult/" + graph;
  mkdirs(qdir);
  const std::string rfile = qdir + "/random.csv";
  const bool exists = file_exists(rfile);
  std::ofstream q(rfile, std::ios::app);
  if (!exists)
    q << "queryNumber,sequence,multiThread-4,multiThread-6,multiThread-8,multiThread-10,multiThread-12,"
         "multiThread-14,multiThread-16,GPUnoLCA-64,GPUwithLCA-64,GPUnoLCA-128,GPUwithLCA-128,GPUnoLCA-256,"
         "GPUwithLCA-256,GPUnoLCA-512,GPUwithLCA-512,GPUnoLCA-1024,GPUwithLCA-1024,TreeHeight,changeHeight,"
         "PartitionMethod,contractDevice\n";
  // CPU columns are not run here (-1); every GPU column holds the plain
  // one-lane-per-query kernel's device time (its answers do not depend on
  // the block size)
  q << nq << ",-1,-1,-1,-1,-1,-1,-1,-1";
  for (int b = 0; b < 5; b++) q << "," << (long long)t_nolca << "," << (long long)t_lca;
  q << "," << info.partition_height << "," << C << "," << method << "," << contract_device << "\n";
  q.close();
  const int rc = graph == "NY" ? real_queries(ctx, info, C, label_kind) : 0;
  h2h_destroy(ctx);
  return rc;
}
