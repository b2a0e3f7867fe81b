// Shortest-path retrieval over the resident index.  The reference's hub
// labels (H2HLabel::Hub, Graph_D.cuh:57-84, set in makeH2HLabel_D,
// kernel_functions.cu:556-676) record which bag member gave each label
// entry; they stop short of unpacking a path into graph edges.  Here a path
// comes from the distances themselves: a neighbour n of the current vertex
// c lies on a shortest c-t path exactly when w(c, n) + d(n, t) = d(c, t), so
// the walk from s needs one distance query per neighbour per hop and no
// index memory beyond the labels.
//
// The device walks every pair at once, one wavefront per pair
// (path_walk_kernel): the next vertex is the positive-weight neighbour of
// least (d(n, t), adjacency position) on a shortest path, so d strictly
// falls and the walk ends at t.  Two passes: the lengths, then the vertices
// into exact offsets.  A pair whose shortest continuations all use
// zero-weight edges is walked on the host instead, as a depth-first search
// that never revisits a vertex and tries candidates in the same order (so a
// strict descent is still taken whenever one exists, and a step back
// happens only on a zero-weight plateau); its rounds answer the queries of
// every such pair in one batched launch.
#include <algorithm>
#include <unordered_set>

#include "h2h_internal.h"

using namespace h2h;

#define H2H_TRY(x)               \
  do {                           \
    h2h_status s_ = (x);         \
    if (s_ != H2H_OK) return s_; \
  } while (0)

namespace {

struct Level {
  int32_t v;
  int64_t d;                       // d(v, t)
  std::vector<int32_t> cand;       // next vertices on a shortest path, in the order tried
  size_t next = 0;
  bool expanded = false;
};

struct Walk {
  int32_t s, t;
  std::vector<Level> stack;
  std::unordered_set<int32_t> seen;
  bool done = false;
};

// The host walk for the pairs `which`: (vertex lists) in that order.
h2h_status host_walks(h2h_ctx* ctx, int kind, const int32_t* s, const int32_t* t, const std::vector<int64_t>& which,
                      std::vector<std::vector<int32_t>>& out) {
  const int n = ctx->n;
  const int64_t m = (int64_t)which.size();
  std::vector<Walk> walks(m);
  std::vector<int32_t> qx, qy, qd;
  for (int64_t i = 0; i < m; i++) {
    qx.push_back(s[which[i]]);
    qy.push_back(t[which[i]]);
  }
  qd.resize(std::max<int64_t>(m, 1));
  h2h_status st = h2h_query(ctx, kind, H2H_EXEC_BATCHED, qx.data(), qy.data(), qd.data(), m, nullptr);
  if (st) return st;
  for (int64_t i = 0; i < m; i++) {
    Walk& w = walks[i];
    w.s = qx[i];
    w.t = qy[i];
    if (qd[i] == INF32) return set_err(ctx, H2H_ERR_INPUT, "no path between a pair (disconnected graph)");
    w.stack.push_back(Level{w.s, qd[i]});
    w.seen.insert(w.s);
    w.done = (w.s == w.t);
  }
  std::vector<std::pair<int64_t, int64_t>> owner;  // (walk, first query) of each expanded level
  for (int64_t round = 0;; round++) {
    if (round > 2 * (int64_t)n + 2)
      return set_err(ctx, H2H_ERR_INPUT, "path retrieval did not finish (labels inconsistent with the graph)");
    // queries for the neighbours of every path's top vertex that is new
    qx.clear();
    qy.clear();
    owner.clear();
    bool active = false;
    for (int64_t i = 0; i < m; i++) {
      Walk& w = walks[i];
      if (w.done) continue;
      active = true;
      if (w.stack.back().expanded) continue;
      const int32_t c = w.stack.back().v;
      owner.emplace_back(i, (int64_t)qx.size());
      for (int64_t e = ctx->adj_off[c]; e < ctx->adj_off[c + 1]; e++) {
        qx.push_back(ctx->adj_nbr[e]);
        qy.push_back(w.t);
      }
    }
    if (!active) break;
    qd.resize(std::max<size_t>(qx.size(), 1));
    if ((st = h2h_query(ctx, kind, H2H_EXEC_BATCHED, qx.data(), qy.data(), qd.data(), (int64_t)qx.size(), nullptr)))
      return st;
    for (const auto& ow : owner) {
      Level& L = walks[ow.first].stack.back();
      const int32_t c = L.v;
      std::vector<std::pair<int64_t, int32_t>> ok;  // (d(n, t), neighbour), adjacency order kept for ties
      int64_t q = ow.second;
      for (int64_t e = ctx->adj_off[c]; e < ctx->adj_off[c + 1]; e++, q++)
        if (qd[q] != INF32 && (int64_t)ctx->adj_w[e] + qd[q] == L.d) ok.emplace_back(qd[q], ctx->adj_nbr[e]);
      std::stable_sort(ok.begin(), ok.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      for (const auto& x : ok) L.cand.push_back(x.second);
      L.expanded = true;
    }
    // advance: push the first unvisited candidate, or step back
    for (int64_t i = 0; i < m; i++) {
      Walk& w = walks[i];
      while (!w.done && w.stack.back().expanded) {
        Level& L = w.stack.back();
        int32_t nx = -1;
        while (L.next < L.cand.size()) {
          const int32_t c = L.cand[L.next++];
          if (!w.seen.count(c)) {
            nx = c;
            break;
          }
        }
        if (nx < 0) {
          w.stack.pop_back();
          if (w.stack.empty())
            return set_err(ctx, H2H_ERR_INPUT, "no shortest path found (labels inconsistent with the graph)");
          continue;
        }
        // the chosen neighbour's distance: d(c, t) - w(c, nx)
        const int32_t c = L.v;
        int64_t wc = 0;
        for (int64_t e = ctx->adj_off[c]; e < ctx->adj_off[c + 1]; e++)
          if (ctx->adj_nbr[e] == nx) {
            wc = ctx->adj_w[e];
            break;
          }
        w.seen.insert(nx);
        w.stack.push_back(Level{nx, L.d - wc});
        if (nx == w.t) w.done = true;
        break;
      }
    }
  }
  out.resize(m);
  for (int64_t i = 0; i < m; i++) {
    out[i].clear();
    for (const Level& L : walks[i].stack) out[i].push_back(L.v);
  }
  return H2H_OK;
}

}  // namespace

extern "C" {

static h2h_status shortest_paths_body(h2h_ctx* ctx, const int32_t* s, const int32_t* t, int64_t npairs,
                              int64_t* total_vertices, double* usec) {
  if (!ctx) return H2H_ERR_ARG;
  if (ctx->device < 0) return set_err(ctx, H2H_ERR_NOGPU, "host-only context (created with device -1)");
  if (!ctx->labels_ready) return set_err(ctx, H2H_ERR_STATE, "build labels first");
  if (!ctx->graph_loaded || (int64_t)ctx->adj_off.size() != (int64_t)ctx->n + 1)
    return set_err(ctx, H2H_ERR_STATE, "path retrieval needs the graph of the index (set_graph or load_files)");
  if (npairs < 0 || (npairs > 0 && (!s || !t))) return set_err(ctx, H2H_ERR_ARG, "bad path pair buffers");
  const int n = ctx->n;
  for (int64_t i = 0; i < npairs; i++)
    if (s[i] < 0 || s[i] >= n || t[i] < 0 || t[i] >= n) return set_err(ctx, H2H_ERR_ARG, "path vertex id out of range");
  (void)hipSetDevice(ctx->device);
  const auto t0 = std::chrono::steady_clock::now();
  const int kind = ctx->label_kind == H2H_LABEL_L4 ? H2H_QUERY_LCA : H2H_QUERY_NOLCA;
  ctx->path_off.assign(npairs + 1, 0);
  ctx->path_vertices.clear();
  if (npairs > 0) {
    // d(s, t), then the device walk in two passes (lengths, vertices)
    std::vector<int32_t> d0(npairs), len(npairs);
    h2h_status st = h2h_query(ctx, kind, H2H_EXEC_BATCHED, s, t, d0.data(), npairs, nullptr);
    if (st) return st;
    for (int64_t i = 0; i < npairs; i++)
      if (d0[i] == INF32) return set_err(ctx, H2H_ERR_INPUT, "no path between a pair (disconnected graph)");
    hipStream_t hs = ctx->stream;
    DevArray d_off, d_nbr, d_w, d_s, d_t, d_d0, d_len, d_poff, d_verts;
    H2H_TRY(upload(ctx, d_off, ctx->adj_off.data(), (n + 1) * sizeof(int64_t)));
    const int32_t pad[1] = {0};  // a graph without arcs (one vertex) still gets valid buffers
    const bool noarc = ctx->adj_nbr.empty();
    H2H_TRY(upload(ctx, d_nbr, noarc ? pad : ctx->adj_nbr.data(), std::max<size_t>(ctx->adj_nbr.size(), 1) * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_w, noarc ? pad : ctx->adj_w.data(), std::max<size_t>(ctx->adj_w.size(), 1) * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_s, s, npairs * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_t, t, npairs * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_d0, d0.data(), npairs * sizeof(int32_t)));
    H2H_TRY(dev_alloc(ctx, d_len, npairs * sizeof(int32_t)));
    H2H_TRY(launch_path_walk(ctx, kind, d_off.as<int64_t>(), d_nbr.as<int32_t>(), d_w.as<int32_t>(), d_s.as<int32_t>(),
                             d_t.as<int32_t>(), d_d0.as<int32_t>(), nullptr, nullptr, d_len.as<int32_t>(), npairs, hs));
    H2H_HIP(ctx, hipMemcpyAsync(len.data(), d_len.p, npairs * sizeof(int32_t), hipMemcpyDeviceToHost, hs));
    H2H_HIP(ctx, hipStreamSynchronize(hs));
    std::vector<int64_t> dev_off(npairs, 0), host_pairs;
    int64_t dev_total = 0;
    for (int64_t i = 0; i < npairs; i++) {
      dev_off[i] = dev_total;
      if (len[i] > 0) dev_total += len[i];
      else host_pairs.push_back(i);
    }
    std::vector<int32_t> dev_verts(std::max<int64_t>(dev_total, 1));
    if (dev_total > 0) {
      H2H_TRY(upload(ctx, d_poff, dev_off.data(), npairs * sizeof(int64_t)));
      H2H_TRY(dev_alloc(ctx, d_verts, dev_total * sizeof(int32_t)));
      H2H_TRY(launch_path_walk(ctx, kind, d_off.as<int64_t>(), d_nbr.as<int32_t>(), d_w.as<int32_t>(),
                               d_s.as<int32_t>(), d_t.as<int32_t>(), d_d0.as<int32_t>(), d_poff.as<int64_t>(),
                               d_verts.as<int32_t>(), d_len.as<int32_t>(), npairs, hs));
      H2H_HIP(ctx, hipMemcpyAsync(dev_verts.data(), d_verts.p, dev_total * sizeof(int32_t), hipMemcpyDeviceToHost, hs));
      H2H_HIP(ctx, hipStreamSynchronize(hs));
    }
    // pairs whose walk met a zero-weight plateau: the host walk
    std::vector<std::vector<int32_t>> hp;
    if (!host_pairs.empty()) H2H_TRY(host_walks(ctx, kind, s, t, host_pairs, hp));
    for (int64_t i = 0, h = 0; i < npairs; i++)
      ctx->path_off[i + 1] = ctx->path_off[i] + (len[i] > 0 ? len[i] : (int64_t)hp[h++].size());
    ctx->path_vertices.resize(ctx->path_off[npairs]);
    for (int64_t i = 0, h = 0; i < npairs; i++) {
      if (len[i] > 0)
        std::copy(dev_verts.begin() + dev_off[i], dev_verts.begin() + dev_off[i] + len[i],
                  ctx->path_vertices.begin() + ctx->path_off[i]);
      else
        std::copy(hp[h].begin(), hp[h].end(), ctx->path_vertices.begin() + ctx->path_off[i]), h++;
    }
  }
  if (total_vertices) *total_vertices = ctx->path_off[npairs];
  if (usec) *usec = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  return H2H_OK;
}

h2h_status h2h_shortest_paths(h2h_ctx* ctx, const int32_t* s, const int32_t* t, int64_t npairs,
                              int64_t* total_vertices, double* usec) {
  return guarded(ctx, [&] { return shortest_paths_body(ctx, s, t, npairs, total_vertices, usec); });
}

h2h_status h2h_get_paths(h2h_ctx* ctx, int64_t* path_off, int32_t* vertices) {
  if (!ctx) return H2H_ERR_ARG;
  if (ctx->path_off.empty()) return set_err(ctx, H2H_ERR_STATE, "no paths computed (h2h_shortest_paths first)");
  if (path_off) std::copy(ctx->path_off.begin(), ctx->path_off.end(), path_off);
  if (vertices) std::copy(ctx->path_vertices.begin(), ctx->path_vertices.end(), vertices);
  return H2H_OK;
}

}  // extern "C"
