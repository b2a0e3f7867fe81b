// Shortest-path retrieval over the resident index.  The reference's hub
// labels (H2HLabel::Hub, Graph_D.cuh:57-84, set in makeH2HLabel_D,
// kernel_functions.cu:556-676) record which bag member gave each label
// entry; they stop short of unpacking a path into graph edges.  Here a path
// comes from the distances themselves: a neighbour n of the current vertex
// c lies on a shortest c-t path exactly when w(c, n) + d(n, t) = d(c, t), so
// the walk from s needs one distance query per neighbour per hop and no
// index memory beyond the labels.  All pairs of a call walk in lockstep:
// each round answers the queries of every path that advanced, in one
// batched launch.
//
// Each path is a depth-first search over such neighbours that never
// revisits a vertex; candidates are taken in order of (d(n, t), adjacency
// order), so a strict descent is taken whenever one exists and a step back
// happens only on a plateau of zero-weight edges.
#include <algorithm>
#include <unordered_set>

#include "h2h_internal.h"

using namespace h2h;

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

}  // namespace

extern "C" {

h2h_status h2h_shortest_paths(h2h_ctx* ctx, const int32_t* s, const int32_t* t, int64_t npairs,
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
  const auto t0 = std::chrono::steady_clock::now();
  const int kind = ctx->label_kind == H2H_LABEL_L4 ? H2H_QUERY_LCA : H2H_QUERY_NOLCA;
  std::vector<Walk> walks(npairs);
  std::vector<int32_t> qx, qy, qd;
  // d(s, t) for every pair
  qx.assign(s, s + npairs);
  qy.assign(t, t + npairs);
  qd.resize(std::max<int64_t>(npairs, 1));
  h2h_status st = h2h_query(ctx, kind, H2H_EXEC_BATCHED, qx.data(), qy.data(), qd.data(), npairs, nullptr);
  if (st) return st;
  for (int64_t i = 0; i < npairs; i++) {
    Walk& w = walks[i];
    w.s = s[i];
    w.t = t[i];
    if (qd[i] == INF32) return set_err(ctx, H2H_ERR_INPUT, "no path between a pair (disconnected graph)");
    w.stack.push_back(Level{s[i], qd[i]});
    w.seen.insert(s[i]);
    w.done = (s[i] == t[i]);
  }
  std::vector<std::pair<int64_t, int64_t>> owner;  // (walk, first query) of each expanded level
  for (int64_t round = 0;; round++) {
    if (round > (int64_t)n + 1) return set_err(ctx, H2H_ERR_INPUT, "path retrieval did not finish (labels inconsistent with the graph)");
    // queries for the neighbours of every path's top vertex that is new
    qx.clear();
    qy.clear();
    owner.clear();
    for (int64_t i = 0; i < npairs; i++) {
      Walk& w = walks[i];
      if (w.done || w.stack.back().expanded) continue;
      const int32_t c = w.stack.back().v;
      owner.emplace_back(i, (int64_t)qx.size());
      for (int64_t e = ctx->adj_off[c]; e < ctx->adj_off[c + 1]; e++) {
        qx.push_back(ctx->adj_nbr[e]);
        qy.push_back(w.t);
      }
    }
    bool active = false;
    for (const Walk& w : walks) active |= !w.done;
    if (!active) break;
    qd.resize(std::max<size_t>(qx.size(), 1));
    if ((st = h2h_query(ctx, kind, H2H_EXEC_BATCHED, qx.data(), qy.data(), qd.data(), (int64_t)qx.size(), nullptr)))
      return st;
    for (const auto& ow : owner) {
      Level& L = walks[ow.first].stack.back();
      const int32_t c = L.v;
      std::vector<std::pair<int64_t, int32_t>> ok;  // (d(n, t), neighbour), adjacency order kept for ties
      int64_t q = ow.second;
      for (int64_t e = ctx->adj_off[c]; e < ctx->adj_off[c + 1]; e++, q++) {
        const int64_t dn = qd[q];
        if (qd[q] != INF32 && (int64_t)ctx->adj_w[e] + dn == L.d) ok.emplace_back(dn, ctx->adj_nbr[e]);
      }
      std::stable_sort(ok.begin(), ok.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      for (const auto& x : ok) L.cand.push_back(x.second);
      L.expanded = true;
    }
    // advance: push the first unvisited candidate, or step back
    for (int64_t i = 0; i < npairs; i++) {
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
          if (w.stack.empty()) return set_err(ctx, H2H_ERR_INPUT, "no shortest path found (labels inconsistent with the graph)");
          continue;
        }
        // the chosen neighbour's distance: d(c, t) - w(c, nx)
        const int32_t c = L.v;
        int64_t wc = -1;
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
  // the paths, as vertex lists, kept on the context for h2h_get_paths
  ctx->path_off.assign(npairs + 1, 0);
  for (int64_t i = 0; i < npairs; i++) ctx->path_off[i + 1] = ctx->path_off[i] + (int64_t)walks[i].stack.size();
  ctx->path_vertices.resize(ctx->path_off[npairs]);
  for (int64_t i = 0; i < npairs; i++) {
    int64_t k = ctx->path_off[i];
    for (const Level& L : walks[i].stack) ctx->path_vertices[k++] = L.v;
  }
  if (total_vertices) *total_vertices = ctx->path_off[npairs];
  if (usec) *usec = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  return H2H_OK;
}

h2h_status h2h_get_paths(h2h_ctx* ctx, int64_t* path_off, int32_t* vertices) {
  if (!ctx) return H2H_ERR_ARG;
  if (ctx->path_off.empty()) return set_err(ctx, H2H_ERR_STATE, "no paths computed (h2h_shortest_paths first)");
  if (path_off) std::copy(ctx->path_off.begin(), ctx->path_off.end(), path_off);
  if (vertices) std::copy(ctx->path_vertices.begin(), ctx->path_vertices.end(), vertices);
  return H2H_OK;
}

}  // extern "C"
