// Distance queries over the resident index.
//
// LCA: preorder range-minimum.  For x != y with tin[x] < tin[y], every vertex
// of the preorder range (tin[x], tin[y]] lies below LCA(x, y) and the
// shallowest of them is a child of it, so the sparse table stores
// (depth(w) << 32 | father(w)) for w = preorder[i] and one min over two
// overlapping power-of-two windows gives the LCA and its depth.  floor(log2)
// is 31 - clz (the reference uses FP64 log2, kernel_functions.cu:1169).
//
// Distances (Query.cu:108-151; kernels kernel_functions.cu:1143-1190, 1253-1296):
//   H2H_QUERY_LCA   min over i in pos(LCA) of dis[x][pos_i] + dis[y][pos_i]
//   H2H_QUERY_NOLCA min over i < depth(LCA)  of dis[x][i]     + dis[y][i]
//
// Execution
//   plain    one lane per query, like the reference kernels;
//   (an id outside [0, n) answers -1 instead of reading out of bounds)
//   batched  queries are first sized (LCA and its width), then split into
//            width classes that run with 4, 16 or 64 lanes per query: a
//            query's label reads become coalesced runs and the lanes of a
//            wavefront finish together.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "h2h_internal.h"

namespace h2h {
namespace {

struct QArgs {
  const int32_t* tin;
  const int64_t* lca;       // [levels][n]
  const int64_t* dis_hash;
  const int64_t* pos_hash;
  const int32_t* pos_pos;
  const int32_t* dis;
  int n;
};

struct Lca {
  int32_t node;
  int32_t depth;
};

__device__ inline Lca find_lca(const QArgs& q, int a, int b) {
  int ta = q.tin[a], tb = q.tin[b];
  if (ta > tb) {
    const int t = ta;
    ta = tb;
    tb = t;
  }
  const int len = tb - ta;  // > 0
  const int k = 31 - __clz(len);
  const int64_t* row = q.lca + (int64_t)k * q.n;
  const int64_t m1 = row[ta + 1];
  const int64_t m2 = row[tb - (1 << k) + 1];
  const int64_t m = m1 < m2 ? m1 : m2;
  Lca r;
  r.node = (int32_t)(m & 0xffffffff);
  r.depth = (int32_t)(m >> 32) - 1;
  return r;
}

__global__ void __launch_bounds__(256) query_plain_kernel(QArgs q, const int32_t* __restrict__ x,
                                                        const int32_t* __restrict__ y, int32_t* __restrict__ out,
                                                        int64_t nq, int kind) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const int a = x[i], b = y[i];
  if ((unsigned)a >= (unsigned)q.n || (unsigned)b >= (unsigned)q.n) {
    out[i] = -1;  // id out of range
    return;
  }
  if (a == b) {
    out[i] = 0;
    return;
  }
  const Lca L = find_lca(q, a, b);
  const int32_t* dx = q.dis + q.dis_hash[a];
  const int32_t* dy = q.dis + q.dis_hash[b];
  int32_t r = INF32;
  if (kind == H2H_QUERY_LCA) {
    const int64_t e1 = q.pos_hash[L.node + 1];
    for (int64_t e = q.pos_hash[L.node]; e < e1; e++) {
      const int p = q.pos_pos[e];
      r = min(r, dx[p] + dy[p]);
    }
  } else {
    for (int j = 0; j < L.depth; j++) r = min(r, dx[j] + dy[j]);
  }
  out[i] = r;
}

// ---- batched ----
// pass 1: LCA, the width of the scan, and the width class of every query
__global__ void __launch_bounds__(256) query_size_kernel(QArgs q, const int32_t* __restrict__ x,
                                                       const int32_t* __restrict__ y, int64_t nq, int kind,
                                                       int32_t* __restrict__ lca_out, int32_t* __restrict__ cls_out,
                                                       int32_t* __restrict__ cls_count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;
  if (i < nq) {
    const int a = x[i], b = y[i];
    int node = -1, width = 0;
    if ((unsigned)a >= (unsigned)q.n || (unsigned)b >= (unsigned)q.n) {
      node = -2;  // id out of range
    } else if (a != b) {
      const Lca L = find_lca(q, a, b);
      node = L.node;
      width = (kind == H2H_QUERY_LCA) ? (int)(q.pos_hash[L.node + 1] - q.pos_hash[L.node]) : L.depth;
    }
    c = width <= 8 ? 0 : (width <= 48 ? 1 : 2);
    lca_out[i] = node;
  }
  // per-wave aggregation of the class counters
  for (int k = 0; k < 3; k++) {
    const uint64_t m = __ballot(c == k);
    if (m && (threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) atomicAdd(&cls_count[k], __popcll(m));
  }
  if (i < nq) cls_out[i] = c;
}

// pass 2: scatter query indices into their class lists
__global__ void __launch_bounds__(256) query_scatter_kernel(const int32_t* __restrict__ cls, int64_t nq,
                                                          const int32_t* __restrict__ cls_base,
                                                          int32_t* __restrict__ cls_fill, int32_t* __restrict__ order) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (i < nq) ? cls[i] : -1;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int k = 0; k < 3; k++) {
    const uint64_t m = __ballot(c == k);
    if (!m) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int basepos = 0;
    if (lane == leader) basepos = atomicAdd(&cls_fill[k], __popcll(m));
    basepos = __shfl(basepos, leader, 64);
    if (c == k) order[cls_base[k] + basepos + __popcll(m & lt)] = (int32_t)i;
  }
}

// pass 3: one persistent launch over the three class lists, widest first;
// class c runs G_c lanes per query, blocks stride over units of 256/G_c queries.
template <int G>
__device__ inline void group_pass(const QArgs& q, const int32_t* __restrict__ x, const int32_t* __restrict__ y,
                                  int32_t* __restrict__ out, const int32_t* __restrict__ lca_node,
                                  const int32_t* __restrict__ order, int64_t cnt, int kind) {
  constexpr int PER = 256 / G;
  const int g = threadIdx.x % G;
  const int64_t units = (cnt + PER - 1) / PER;
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const int64_t slot = u * PER + threadIdx.x / G;
    int32_t r = INF32;
    int64_t qi = -1;
    if (slot < cnt) {
      qi = order[slot];
      const int L = lca_node[qi];
      if (L < 0) {
        r = (L == -1) ? 0 : -1;
      } else {
        const int32_t* dx = q.dis + q.dis_hash[x[qi]];
        const int32_t* dy = q.dis + q.dis_hash[y[qi]];
        // four strides per round: the position loads, then the eight label
        // loads, are in flight together (clamped indices stay in the rows)
        if (kind == H2H_QUERY_LCA) {
          const int64_t e0 = q.pos_hash[L], e1 = q.pos_hash[L + 1];
          for (int64_t e = e0 + g; e < e1; e += 4 * G) {
            int p[4];
#pragma unroll
            for (int k = 0; k < 4; k++) p[k] = q.pos_pos[min(e + k * G, e1 - 1)];
            int32_t a[4], b[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
              a[k] = dx[p[k]];
              b[k] = dy[p[k]];
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
              if (e + k * G < e1) r = min(r, a[k] + b[k]);
          }
        } else {
          const int depthL = (int)(q.dis_hash[L + 1] - q.dis_hash[L]);
          for (int j = g; j < depthL; j += 4 * G) {
            int32_t a[4], b[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int jj = min(j + k * G, depthL - 1);
              a[k] = dx[jj];
              b[k] = dy[jj];
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
              if (j + k * G < depthL) r = min(r, a[k] + b[k]);
          }
        }
      }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) r = min(r, __shfl_xor(r, off, G));
    if (qi >= 0 && g == 0) out[qi] = r;
  }
}

__global__ void __launch_bounds__(256) query_group_kernel(QArgs q, const int32_t* __restrict__ x,
                                                        const int32_t* __restrict__ y, int32_t* __restrict__ out,
                                                        const int32_t* __restrict__ lca_node,
                                                        const int32_t* __restrict__ order,
                                                        const int32_t* __restrict__ cls_count, int kind) {
  const int64_t c0 = cls_count[0], c1 = cls_count[1], c2 = cls_count[2];
  group_pass<64>(q, x, y, out, lca_node, order + c0 + c1, c2, kind);
  group_pass<16>(q, x, y, out, lca_node, order + c0, c1, kind);
  group_pass<4>(q, x, y, out, lca_node, order, c0, kind);
}

// class bookkeeping: zero the counters / turn counts into list bases
__global__ void query_zero_kernel(int32_t* v, int n) {
  if ((int)threadIdx.x < n) v[threadIdx.x] = 0;
}
__global__ void query_bases_kernel(const int32_t* cls_count, int32_t* cls_base, int32_t* cls_fill) {
  if (threadIdx.x == 0) {
    cls_base[0] = 0;
    cls_base[1] = cls_count[0];
    cls_base[2] = cls_count[0] + cls_count[1];
    cls_fill[0] = cls_fill[1] = cls_fill[2] = 0;
  }
}

}  // namespace

#define H2H_TRY(x)               \
  do {                           \
    h2h_status s_ = (x);         \
    if (s_ != H2H_OK) return s_; \
  } while (0)

h2h_status build_lca(h2h_ctx* ctx) {
  const int n = ctx->n;
  int K = 1;
  while ((1LL << K) <= n) K++;
  std::vector<int64_t> t((size_t)K * n, 0);
  for (int i = 0; i < n; i++) {
    const int w = ctx->preorder[i];
    t[i] = ((int64_t)ctx->depth[w] << 32) | (uint32_t)ctx->father[w];
  }
  for (int k = 1; k < K; k++) {
    const int half = 1 << (k - 1);
    int64_t* cur = t.data() + (size_t)k * n;
    const int64_t* prev = t.data() + (size_t)(k - 1) * n;
    for (int i = 0; i + (1 << k) <= n; i++) cur[i] = std::min(prev[i], prev[i + half]);
  }
  ctx->lca_levels = K;
  H2H_TRY(upload(ctx, ctx->d_lca, t.data(), t.size() * sizeof(int64_t)));
  H2H_TRY(upload(ctx, ctx->d_tin, ctx->tin.data(), n * sizeof(int32_t)));
  return H2H_OK;
}

h2h_status launch_query(h2h_ctx* ctx, int kind, int exec, const int32_t* dx, const int32_t* dy, int32_t* dout,
                        int64_t nq, hipStream_t s, QueryWorkspace* ws) {
  if (nq <= 0) return H2H_OK;
  QArgs q;
  q.tin = ctx->d_tin.as<int32_t>();
  q.lca = ctx->d_lca.as<int64_t>();
  q.dis_hash = ctx->d_dis_hash.as<int64_t>();
  q.pos_hash = ctx->d_pos_hash.as<int64_t>();
  q.pos_pos = ctx->d_pos_pos.as<int32_t>();
  q.dis = ctx->d_dis.as<int32_t>();
  q.n = ctx->n;
  const unsigned blocks = (unsigned)((nq + 255) / 256);
  if (exec == H2H_EXEC_PLAIN) {
    hipLaunchKernelGGL(query_plain_kernel, dim3(blocks), dim3(256), 0, s, q, dx, dy, dout, nq, kind);
    H2H_HIP(ctx, hipGetLastError());
    return H2H_OK;
  }
  if (!ws || ws->cap < nq) return set_err(ctx, H2H_ERR_ARG, "batched query workspace too small");
  int32_t* lca_node = ws->lca.as<int32_t>();
  int32_t* cls = ws->cls.as<int32_t>();
  int32_t* order = ws->order.as<int32_t>();
  int32_t* ctr = ws->ctr.as<int32_t>();  // [0,3) counts, [3,6) bases, [6,9) fill
  hipLaunchKernelGGL(query_zero_kernel, dim3(1), dim3(64), 0, s, ctr, 9);
  hipLaunchKernelGGL(query_size_kernel, dim3(blocks), dim3(256), 0, s, q, dx, dy, nq, kind, lca_node, cls, ctr);
  hipLaunchKernelGGL(query_bases_kernel, dim3(1), dim3(64), 0, s, ctr, ctr + 3, ctr + 6);
  hipLaunchKernelGGL(query_scatter_kernel, dim3(blocks), dim3(256), 0, s, cls, nq, ctr + 3, ctr + 6, order);
  unsigned grid = (unsigned)std::min<int64_t>((nq + 3) / 4, (int64_t)ctx->num_cus * 8);
  if (grid == 0) grid = 1;
  hipLaunchKernelGGL(query_group_kernel, dim3(grid), dim3(256), 0, s, q, dx, dy, dout, lca_node, order, ctr, kind);
  H2H_HIP(ctx, hipGetLastError());
  return H2H_OK;
}

h2h_status query_workspace(h2h_ctx* ctx, QueryWorkspace& ws, int64_t nq) {
  if (ws.cap >= nq) return H2H_OK;
  const int64_t cap = std::max<int64_t>(nq, 1024);
  H2H_TRY(dev_alloc(ctx, ws.lca, cap * sizeof(int32_t)));
  H2H_TRY(dev_alloc(ctx, ws.cls, cap * sizeof(int32_t)));
  H2H_TRY(dev_alloc(ctx, ws.order, cap * sizeof(int32_t)));
  H2H_TRY(dev_alloc(ctx, ws.ctr, 16 * sizeof(int32_t)));
  ws.cap = cap;
  return H2H_OK;
}

}  // namespace h2h
