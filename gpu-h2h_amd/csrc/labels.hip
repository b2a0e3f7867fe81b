// Label construction on the GPU and the resident index.
//
// Layout in HBM (the reference layout, H2HConsturction.cu:618-722):
//   dis[dis_hash[v] + j], j < depth(v): distance from v to its ancestor at
//     depth j+1 (j = depth(v)-1 is v itself, 0);
//   pos_pos/pos_w/pos_rowoff[pos_hash[v] ..]: the bag X(v) sorted by position
//     p_u = depth(u)-1 with the shortcut weight w(v,u) and dis_hash[u], then
//     v itself last.
//
// A node's row only depends on its ancestors' rows, so the tree is built
// top-down level by level.  For L3 (as shipped) and L4 (exact) the row is
//   dis[v][j] = min over u in X(v) of  w(v,u) + (p_u >= j ? dis[u][j]
//                                                   : L4 ? dis[a_j][p_u] : inf)
// where a_j is v's ancestor at depth j+1.  This is the fixpoint the
// reference's deepest-first loop reaches (H2HConsturction.cu:766-790 with the
// L4 loop of :783-789): the labels satisfy the triangle inequality, so the
// update order and the use of w(v,u) instead of the running dis[v][p_u] do
// not change any entry.
//
// Kernels
//   label_pf_kernel   one workgroup per node of a level (parallelFine = 1):
//                     lanes own columns j (coalesced reads of ancestor rows),
//                     the workgroup's waves split the bag entries and combine
//                     through LDS.
//   label_ll_*        one wavefront per (node, bag entry) of a level with an
//                     atomic min into the node's row (parallelFine = 2, the
//                     rule of kernel_functions.cu:717-744).
//   label_dataflow    one persistent launch over the whole tree: workgroups
//                     take nodes in BFS order from a counter and start a node
//                     as soon as its father's row is published (no level
//                     barriers), method H2H_METHOD_BATCHED.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "h2h_internal.h"

namespace h2h {
namespace {

constexpr int R = 8;  // columns per lane per pass: a pass covers 512 columns

struct LabelArgs {
  const int32_t* bfs;
  const int32_t* father;
  const int64_t* dis_hash;
  const int64_t* pos_hash;
  const int32_t* pos_pos;
  const int32_t* pos_w;
  const int64_t* pos_rowoff;
  int32_t* dis;
  int32_t* anc;  // ancestor ids, same layout as dis (L4 only)
  int l4;
};

// Computes columns [0, L) of node v's row with `nw` waves of the calling
// workgroup (wave index wid).  comb: 64*R ints of LDS.
__device__ inline void node_row(const LabelArgs& a, int v, int d, int wid, int nw, int lane,
                                int32_t* comb) {
  const int64_t rowv = a.dis_hash[v];
  const int64_t e0 = a.pos_hash[v];
  const int64_t e1 = a.pos_hash[v + 1] - 1;  // the self entry is last
  const int L = d - 1;
  int64_t ancf = 0;
  if (d > 1 && a.l4) ancf = a.dis_hash[a.father[v]];
  if (a.l4) {
    for (int t = wid * 64 + lane; t < L; t += nw * 64) a.anc[rowv + t] = a.anc[ancf + t];
    if (wid == 0 && lane == 0) a.anc[rowv + L] = v;
  }
  for (int c0 = 0; c0 < L; c0 += 64 * R) {
    int32_t acc[R];
    int64_t aoff[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      acc[r] = INF32;
      aoff[r] = 0;
      const int j = c0 + r * 64 + lane;
      if (a.l4 && j < L) aoff[r] = a.dis_hash[a.anc[ancf + j]];
    }
    for (int64_t e = e0 + wid; e < e1; e += nw) {
      const int p = a.pos_pos[e];
      const int w = a.pos_w[e];
      const int64_t uo = a.pos_rowoff[e];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        if (j < L) {
          if (j <= p) {
            acc[r] = min(acc[r], w + a.dis[uo + j]);
          } else if (a.l4) {
            acc[r] = min(acc[r], w + a.dis[aoff[r] + p]);
          }
        }
      }
    }
    if (nw == 1) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        if (j < L) a.dis[rowv + j] = acc[r];
      }
    } else {
      for (int t = wid * 64 + lane; t < 64 * R; t += nw * 64) comb[t] = INF32;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        if (j < L && acc[r] != INF32) atomicMin(&comb[r * 64 + lane], acc[r]);
      }
      __syncthreads();
      for (int t = wid * 64 + lane; t < 64 * R; t += nw * 64) {
        const int j = c0 + t;
        if (j < L) a.dis[rowv + j] = comb[t];
      }
      __syncthreads();
    }
  }
  if (wid == 0 && lane == 0) a.dis[rowv + L] = 0;
}

__global__ void __launch_bounds__(1024) label_pf_kernel(LabelArgs a, int lvl_begin, int d) {
  __shared__ int32_t comb[64 * R];
  const int v = a.bfs[lvl_begin + blockIdx.x];
  node_row(a, v, d, threadIdx.x >> 6, blockDim.x >> 6, threadIdx.x & 63, comb);
}

// ---- LL: row initialisation as mallocH2HLabel_noHub leaves it, then one
// wavefront per (node, bag entry) with atomic min (kernel_functions.cu:717-744).
__global__ void __launch_bounds__(256) label_ll_init_kernel(LabelArgs a, int lvl_begin, int count, int d) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= count) return;
  const int v = a.bfs[lvl_begin + wave];
  const int64_t rowv = a.dis_hash[v];
  const int L = d - 1;
  for (int j = lane; j < L; j += 64) a.dis[rowv + j] = INF32;
  if (a.l4 && d > 1) {
    const int64_t ancf = a.dis_hash[a.father[v]];
    for (int j = lane; j < L; j += 64) a.anc[rowv + j] = a.anc[ancf + j];
  }
  // bag weights at the bag positions (positions are distinct)
  for (int64_t e = a.pos_hash[v] + lane; e < a.pos_hash[v + 1] - 1; e += 64) a.dis[rowv + a.pos_pos[e]] = a.pos_w[e];
  if (lane == 0) {
    a.dis[rowv + L] = 0;
    if (a.l4) a.anc[rowv + L] = v;
  }
}

__global__ void __launch_bounds__(256) label_ll_kernel(LabelArgs a, const int32_t* ent_node,
                                                      const int64_t* ent_pos, int64_t begin, int64_t count, int d) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= count) return;
  const int v = ent_node[begin + wave];
  const int64_t e = ent_pos[begin + wave];
  const int p = a.pos_pos[e];
  const int64_t rowv = a.dis_hash[v];
  const int64_t uo = a.pos_rowoff[e];
  const int t = __hip_atomic_load(&a.dis[rowv + p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int j = lane; j < p; j += 64) atomicMin(&a.dis[rowv + j], t + a.dis[uo + j]);
  if (a.l4) {
    for (int j = p + 1 + lane; j < d - 1; j += 64) {
      const int64_t ao = a.dis_hash[a.anc[rowv + j]];
      atomicMin(&a.dis[rowv + j], t + a.dis[ao + p]);
    }
  }
}

// ---- dataflow: one persistent launch; a node starts once its father's row
// is published.  Nodes are taken in BFS order, so every father was taken
// earlier by a running workgroup and the waits end.
struct FlowArgs {
  LabelArgs a;
  const int32_t* bfs_depth;   // depth of bfs[i]
  const int32_t* bfs_index;   // [n] position of v in bfs
  int32_t* done;              // [n] 1 once the row of bfs[i] is published
  int32_t* next;              // work counter
  int32_t* err;               // set if a wait exceeds its spin limit
  int n;
  int spin_limit;
};

// No lane-divergent branch may sit in this loop: with `if (threadIdx.x == 0)`
// around the fetch and the publish, the compiler joins the two across the
// back edge into a region only lane 0 enters, and the other lanes loop back to
// the barrier on the stale item.  Wave 0 therefore fetches, waits and
// publishes as a whole (every lane runs the same atomic / load / store).
__global__ void __launch_bounds__(256) label_dataflow_kernel(FlowArgs f) {
  __shared__ int32_t comb[64 * R];
  __shared__ int32_t s_item;
  __shared__ int32_t s_ok;
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = threadIdx.x & 63;
  for (;;) {
    if (wid == 0) {
      // one increment per block: lane 0 adds 1, the other lanes add 0
      int t = atomicAdd(f.next, lane == 0 ? 1 : 0);
      t = __shfl(t, 0, 64);
      s_item = t;
    }
    __syncthreads();
    const int item = s_item;
    if (item >= f.n) break;
    const int v = f.a.bfs[item];
    const int d = f.bfs_depth[item];
    if (wid == 0) {
      int ok = 1;
      if (d > 1) {
        const int fi = f.bfs_index[f.a.father[v]];
        int spins = 0;
        while (__hip_atomic_load(&f.done[fi], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
          if (__hip_atomic_load(f.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > f.spin_limit) {
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        if (!ok) __hip_atomic_store(f.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      s_ok = ok;
    }
    __syncthreads();
    if (s_ok) node_row(f.a, v, d, wid, nw, lane, comb);
    // every wave's row stores reach device scope before the row is published
    __threadfence();
    __syncthreads();
    if (wid == 0) __hip_atomic_store(&f.done[item], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

int pick_wpn(int64_t nodes, int64_t entries) {
  const double avg = nodes ? (double)entries / nodes : 0.0;
  int w = 1;
  while (w < 16 && w * 8 < avg) w *= 2;
  if (nodes >= 4096) w = std::min(w, 2);
  if (nodes >= 16384) w = 1;
  return w;
}

}  // namespace

#define H2H_TRY(x)               \
  do {                           \
    h2h_status s_ = (x);         \
    if (s_ != H2H_OK) return s_; \
  } while (0)

h2h_status upload_index(h2h_ctx* ctx) {
  const int n = ctx->n;
  const int64_t P = ctx->pos_hash[n];
  std::vector<int64_t> rowoff(P);
  for (int64_t k = 0; k < P; k++) rowoff[k] = ctx->dis_hash[ctx->pos_id[k]];
  H2H_TRY(upload(ctx, ctx->d_dis_hash, ctx->dis_hash.data(), (n + 1) * sizeof(int64_t)));
  H2H_TRY(upload(ctx, ctx->d_pos_hash, ctx->pos_hash.data(), (n + 1) * sizeof(int64_t)));
  H2H_TRY(upload(ctx, ctx->d_pos_pos, ctx->pos_pos.data(), P * sizeof(int32_t)));
  H2H_TRY(upload(ctx, ctx->d_pos_w, ctx->pos_w.data(), P * sizeof(int32_t)));
  H2H_TRY(upload(ctx, ctx->d_pos_rowoff, rowoff.data(), P * sizeof(int64_t)));
  H2H_TRY(upload(ctx, ctx->d_bfs, ctx->bfs.data(), n * sizeof(int32_t)));
  H2H_TRY(upload(ctx, ctx->d_father, ctx->father.data(), n * sizeof(int32_t)));
  H2H_TRY(upload(ctx, ctx->d_depth, ctx->depth.data(), n * sizeof(int32_t)));
  H2H_TRY(dev_alloc(ctx, ctx->d_dis, std::max<int64_t>(ctx->label_entries, 1) * sizeof(int32_t)));
  H2H_TRY(build_lca(ctx));
  H2H_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->labels_ready = false;
  return H2H_OK;
}

h2h_status run_label_build(h2h_ctx* ctx, int kind, int method, double* usec) {
  const int n = ctx->n;
  hipStream_t s = ctx->stream;
  const bool l4 = (kind == H2H_LABEL_L4);
  if (l4) H2H_TRY(dev_alloc(ctx, ctx->d_anc, std::max<int64_t>(ctx->label_entries, 1) * sizeof(int32_t)));
  LabelArgs a;
  a.bfs = ctx->d_bfs.as<int32_t>();
  a.father = ctx->d_father.as<int32_t>();
  a.dis_hash = ctx->d_dis_hash.as<int64_t>();
  a.pos_hash = ctx->d_pos_hash.as<int64_t>();
  a.pos_pos = ctx->d_pos_pos.as<int32_t>();
  a.pos_w = ctx->d_pos_w.as<int32_t>();
  a.pos_rowoff = ctx->d_pos_rowoff.as<int64_t>();
  a.dis = ctx->d_dis.as<int32_t>();
  a.anc = l4 ? ctx->d_anc.as<int32_t>() : nullptr;
  a.l4 = l4 ? 1 : 0;
  const int levels = (int)ctx->level_off.size() - 1;

  // per-level bag entry counts (entries exclude the self entry)
  std::vector<int64_t> lvl_entries(levels, 0);
  for (int l = 0; l < levels; l++)
    for (int i = ctx->level_off[l]; i < ctx->level_off[l + 1]; i++) {
      int v = ctx->bfs[i];
      lvl_entries[l] += ctx->pos_hash[v + 1] - ctx->pos_hash[v] - 1;
    }

  DevArray d_ent_node, d_ent_pos, d_flow_depth, d_flow_index, d_done, d_next, d_err;
  std::vector<int64_t> ent_begin;
  if (method == H2H_METHOD_LL) {
    std::vector<int32_t> en;
    std::vector<int64_t> ep;
    ent_begin.push_back(0);
    for (int l = 0; l < levels; l++) {
      for (int i = ctx->level_off[l]; i < ctx->level_off[l + 1]; i++) {
        int v = ctx->bfs[i];
        for (int64_t e = ctx->pos_hash[v]; e < ctx->pos_hash[v + 1] - 1; e++) {
          en.push_back(v);
          ep.push_back(e);
        }
      }
      ent_begin.push_back((int64_t)en.size());
    }
    H2H_TRY(upload(ctx, d_ent_node, en.data(), en.size() * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_ent_pos, ep.data(), ep.size() * sizeof(int64_t)));
  } else if (method == H2H_METHOD_BATCHED) {
    std::vector<int32_t> bd(n), bi(n);
    for (int l = 0; l < levels; l++)
      for (int i = ctx->level_off[l]; i < ctx->level_off[l + 1]; i++) {
        bd[i] = l + 1;
        bi[ctx->bfs[i]] = i;
      }
    H2H_TRY(upload(ctx, d_flow_depth, bd.data(), n * sizeof(int32_t)));
    H2H_TRY(upload(ctx, d_flow_index, bi.data(), n * sizeof(int32_t)));
    H2H_TRY(dev_alloc(ctx, d_done, n * sizeof(int32_t)));
    H2H_TRY(dev_alloc(ctx, d_next, sizeof(int32_t)));
    H2H_TRY(dev_alloc(ctx, d_err, sizeof(int32_t)));
  }

  hipEvent_t e0, e1;
  H2H_HIP(ctx, hipEventCreate(&e0));
  H2H_HIP(ctx, hipEventCreate(&e1));
  if (method == H2H_METHOD_BATCHED) {
    H2H_HIP(ctx, hipMemsetAsync(d_done.p, 0, n * sizeof(int32_t), s));
    H2H_HIP(ctx, hipMemsetAsync(d_next.p, 0, sizeof(int32_t), s));
    H2H_HIP(ctx, hipMemsetAsync(d_err.p, 0, sizeof(int32_t), s));
  }
  H2H_HIP(ctx, hipEventRecord(e0, s));
  if (method == H2H_METHOD_PF) {
    for (int l = 0; l < levels; l++) {
      const int cnt = ctx->level_off[l + 1] - ctx->level_off[l];
      const int wpn = pick_wpn(cnt, lvl_entries[l]);
      hipLaunchKernelGGL(label_pf_kernel, dim3(cnt), dim3(64 * wpn), 0, s, a, ctx->level_off[l], l + 1);
    }
  } else if (method == H2H_METHOD_LL) {
    for (int l = 0; l < levels; l++) {
      const int cnt = ctx->level_off[l + 1] - ctx->level_off[l];
      hipLaunchKernelGGL(label_ll_init_kernel, dim3((cnt + 3) / 4), dim3(256), 0, s, a, ctx->level_off[l], cnt, l + 1);
      const int64_t ecnt = ent_begin[l + 1] - ent_begin[l];
      if (ecnt > 0)
        hipLaunchKernelGGL(label_ll_kernel, dim3((unsigned)((ecnt + 3) / 4)), dim3(256), 0, s, a,
                           d_ent_node.as<int32_t>(), d_ent_pos.as<int64_t>(), ent_begin[l], ecnt, l + 1);
    }
  } else {
    FlowArgs f;
    f.a = a;
    f.bfs_depth = d_flow_depth.as<int32_t>();
    f.bfs_index = d_flow_index.as<int32_t>();
    f.done = d_done.as<int32_t>();
    f.next = d_next.as<int32_t>();
    f.err = d_err.as<int32_t>();
    f.n = n;
    f.spin_limit = 1 << 20;  // about a second of polling: far beyond any father's row
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) cus = prop.multiProcessorCount;
    hipLaunchKernelGGL(label_dataflow_kernel, dim3(cus * 4), dim3(256), 0, s, f);
  }
  H2H_HIP(ctx, hipGetLastError());
  H2H_HIP(ctx, hipEventRecord(e1, s));
  H2H_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0;
  H2H_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (method == H2H_METHOD_BATCHED) {
    int32_t herr = 0;
    H2H_HIP(ctx, hipMemcpy(&herr, d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (herr) return set_err(ctx, H2H_ERR_HIP, "dataflow label build: a wait exceeded its spin limit");
  }
  ctx->t_label_us = ms * 1000.0;
  if (usec) *usec = ctx->t_label_us;
  ctx->label_kind = kind;
  ctx->labels_ready = true;
  return H2H_OK;
}

}  // namespace h2h
