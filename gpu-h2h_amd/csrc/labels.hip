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
//   label_items_kernel  one workgroup per work item = (node, range of its bag
//                       entries) of one level: lanes own columns j (coalesced
//                       reads of the ancestor rows), waves split the entries and
//                       combine through LDS; items sharing a node combine with
//                       atomicMin.
//   label_init_kernel   INF rows for nodes split over several items.
// Methods (work-item plans over the same kernel)
//   PF       one item per node (parallelFine = 1, kernel_functions.cu:803-842)
//   LL       one item per (node, bag entry) (parallelFine = 2, the atomicMin
//            rule of kernel_functions.cu:717-744)
//   BATCHED  per level, one item per node where the level is wide and bags
//            split into groups where it is narrow; the whole sweep is captured
//            once in a hipGraph and replayed.
// A workgroup that waits for another workgroup's row inside one launch needs
// a device-scope release per row, which on this part writes back the L2; the
// sweep therefore synchronises levels at kernel boundaries instead.
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
  uint32_t* anc;  // row offset of the ancestor at each column, same layout as dis (L4 only)
  int l4;
  int idx32;      // every dis byte offset fits 32 bits
};

// One work item = (node, a contiguous range of its bag entries).  The block's
// waves split the range, each lane owns R columns of a 512-column pass and
// keeps the running minimum in registers; waves combine through LDS.  An item
// that covers the whole bag stores its row; items that share a node combine
// with one atomicMin per column (the row was set to INF beforehand).
// For L4 the column j > p_u gathers dis[a_j][p_u]; a_j (j < depth-1) is read
// from the father's ancestor row, so no item waits for its own node's row.
struct Item {
  int32_t node;   // index into bfs
  int32_t first;  // first bag entry, relative to pos_hash[v]
  int32_t count;  // bag entries of this item
  int32_t flags;  // ITEM_DIRECT: covers the whole bag; ITEM_LEAD: writes self entry / ancestor row
};
constexpr int ITEM_DIRECT = 1;
constexpr int ITEM_LEAD = 2;

// Loads of one bag entry are issued as a batch (addresses first, then all R
// loads, then the mins): a load guarded by a lane-divergent branch makes the
// compiler wait for it before the next one.  With IDX32 every dis index is a
// 32-bit byte offset from the (scalar) base, so each load needs one address
// register instead of two.
template <bool IDX32>
__device__ inline int32_t ld(const int32_t* base, uint64_t idx) {
  if (IDX32) return *(const int32_t*)((const char*)base + (uint32_t)idx * 4u);
  return base[idx];
}

// Accumulates bag entries e0 + k * stride (k < nk) into the columns
// [c0, c0 + 64R) of a row of length L + 1.  The entries' (p, w, row) sit one
// per lane (64 at a time) and are broadcast with readlane, two entries per
// round of loads.
template <bool L4, bool IDX32>
__device__ inline void accumulate(const LabelArgs& a, int64_t e0, int stride, int nk, int c0, int L, int lane,
                                  const uint32_t* aoff, int32_t* acc) {
  for (int kb = 0; kb < nk; kb += 64) {
    int mp = 0, mw = 0;
    uint64_t mu = 0;
    if (kb + lane < nk) {
      const int64_t e = e0 + (int64_t)(kb + lane) * stride;
      mp = a.pos_pos[e];
      mw = a.pos_w[e];
      mu = (uint64_t)a.pos_rowoff[e];
    }
    const int mlo = (int)(uint32_t)mu, mhi = (int)(mu >> 32);
    const int kn = min(64, nk - kb);
    int i = 0;
    for (; i + 1 < kn; i += 2) {
      const int p0 = __builtin_amdgcn_readlane(mp, i), p1 = __builtin_amdgcn_readlane(mp, i + 1);
      const int w0 = __builtin_amdgcn_readlane(mw, i), w1 = __builtin_amdgcn_readlane(mw, i + 1);
      const uint64_t u0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(mhi, i) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane(mlo, i);
      const uint64_t u1 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(mhi, i + 1) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane(mlo, i + 1);
      int32_t v0[R], v1[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        // L3 reads a clamped, always valid address and discards it past p
        v0[r] = ld<IDX32>(a.dis, L4 ? ((j <= p0) ? u0 + j : (uint64_t)aoff[r] + p0) : u0 + min(j, p0));
        v1[r] = ld<IDX32>(a.dis, L4 ? ((j <= p1) ? u1 + j : (uint64_t)aoff[r] + p1) : u1 + min(j, p1));
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        const bool in = j < L;
        acc[r] = (in && (L4 || j <= p0)) ? min(acc[r], w0 + v0[r]) : acc[r];
        acc[r] = (in && (L4 || j <= p1)) ? min(acc[r], w1 + v1[r]) : acc[r];
      }
    }
    if (i < kn) {
      const int p0 = __builtin_amdgcn_readlane(mp, i);
      const int w0 = __builtin_amdgcn_readlane(mw, i);
      const uint64_t u0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(mhi, i) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane(mlo, i);
      int32_t v0[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        v0[r] = ld<IDX32>(a.dis, L4 ? ((j <= p0) ? u0 + j : (uint64_t)aoff[r] + p0) : u0 + min(j, p0));
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        acc[r] = ((j < L) && (L4 || j <= p0)) ? min(acc[r], w0 + v0[r]) : acc[r];
      }
    }
  }
}

template <bool L4, bool IDX32>
__device__ inline void item_row(const LabelArgs& a, int v, int d, int64_t e0, int64_t e1, int flags, int wid,
                                int nw, int lane, int32_t* comb) {
  const int64_t rowv = a.dis_hash[v];
  const int L = d - 1;
  const int nk = (e1 > e0 + wid) ? (int)((e1 - e0 - wid + nw - 1) / nw) : 0;
  int64_t ancf = 0;
  if (L4 && d > 1) ancf = a.dis_hash[a.father[v]];
  if (L4 && (flags & ITEM_LEAD)) {
    for (int t = wid * 64 + lane; t < L; t += nw * 64) a.anc[rowv + t] = a.anc[ancf + t];
    if (wid == 0 && lane == 0) a.anc[rowv + L] = (uint32_t)rowv;
  }
  for (int c0 = 0; c0 < L; c0 += 64 * R) {
    int32_t acc[R];
    uint32_t aoff[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      acc[r] = INF32;
      const int j = c0 + r * 64 + lane;
      aoff[r] = (L4 && j < L) ? a.anc[ancf + j] : 0u;
    }
    accumulate<L4, IDX32>(a, e0 + wid, nw, nk, c0, L, lane, aoff, acc);
    if (nw > 1) {
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
        if (j < L) {
          if (flags & ITEM_DIRECT) a.dis[rowv + j] = comb[t];
          else if (comb[t] != INF32) atomicMin(&a.dis[rowv + j], comb[t]);
        }
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = c0 + r * 64 + lane;
        if (j < L) {
          if (flags & ITEM_DIRECT) a.dis[rowv + j] = acc[r];
          else if (acc[r] != INF32) atomicMin(&a.dis[rowv + j], acc[r]);
        }
      }
    }
  }
  if ((flags & ITEM_LEAD) && wid == 0 && lane == 0) a.dis[rowv + L] = 0;
}

// One wavefront computes the columns [c0, c0 + 64R) of node v's row from its
// whole bag and stores them; the unit with c0 = 0 also writes the self entry.
template <bool L4, bool IDX32>
__device__ inline void unit_row(const LabelArgs& a, int v, int d, int c0, int lane) {
  const int64_t rowv = a.dis_hash[v];
  const int L = d - 1;
  const int64_t e0 = a.pos_hash[v];
  const int nk = (int)(a.pos_hash[v + 1] - 1 - e0);
  int64_t ancf = 0;
  if (L4 && d > 1) ancf = a.dis_hash[a.father[v]];
  if (L4) {
    const int jend = min(c0 + 64 * R, L);
    for (int j = c0 + lane; j < jend; j += 64) a.anc[rowv + j] = a.anc[ancf + j];
    if (c0 == 0 && lane == 0) a.anc[rowv + L] = (uint32_t)rowv;
  }
  if (c0 == 0 && lane == 0) a.dis[rowv + L] = 0;
  if (c0 >= L) return;
  int32_t acc[R];
  uint32_t aoff[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    acc[r] = INF32;
    const int j = c0 + r * 64 + lane;
    aoff[r] = (L4 && j < L) ? a.anc[ancf + j] : 0u;
  }
  accumulate<L4, IDX32>(a, e0, 1, nk, c0, L, lane, aoff, acc);
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int j = c0 + r * 64 + lane;
    if (j < L) a.dis[rowv + j] = acc[r];
  }
}

// Workgroups go to the 8 XCDs round-robin; block b takes item
// (b % 8) * chunk + b / 8 so that each XCD works through a contiguous run of
// the level's items (siblings, which read the same ancestor rows, share an L2).
constexpr int XCDS = 8;

template <bool L4, bool IDX32>
__global__ void __launch_bounds__(256) label_items_kernel(LabelArgs a, const Item* __restrict__ items, int begin,
                                                        int count, int d) {
  __shared__ int32_t comb[64 * R];
  const int chunk = (count + XCDS - 1) / XCDS;
  const int idx = (int)(blockIdx.x % XCDS) * chunk + (int)(blockIdx.x / XCDS);
  if (idx >= count) return;
  const Item it = items[begin + idx];
  const int v = a.bfs[it.node];
  const int64_t e0 = a.pos_hash[v] + it.first;
  item_row<L4, IDX32>(a, v, d, e0, e0 + it.count, it.flags, threadIdx.x >> 6, blockDim.x >> 6, threadIdx.x & 63,
                      comb);
}

void launch_items(const LabelArgs& a, int count, dim3 block, hipStream_t s, const Item* items, int begin, int d) {
  const dim3 grid((unsigned)(((count + XCDS - 1) / XCDS) * XCDS));
  if (a.idx32) {
    if (a.l4) hipLaunchKernelGGL((label_items_kernel<true, true>), grid, block, 0, s, a, items, begin, count, d);
    else hipLaunchKernelGGL((label_items_kernel<false, true>), grid, block, 0, s, a, items, begin, count, d);
  } else {
    if (a.l4) hipLaunchKernelGGL((label_items_kernel<true, false>), grid, block, 0, s, a, items, begin, count, d);
    else hipLaunchKernelGGL((label_items_kernel<false, false>), grid, block, 0, s, a, items, begin, count, d);
  }
}

// Rows of nodes whose items combine with atomicMin start at INF (self 0).
__global__ void __launch_bounds__(256) label_init_kernel(LabelArgs a, const int32_t* __restrict__ nodes,
                                                       const int32_t* __restrict__ depth_of, int count) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= count) return;
  const int v = a.bfs[nodes[wave]];
  const int d = depth_of[wave];
  const int64_t rowv = a.dis_hash[v];
  for (int j = lane; j < d - 1; j += 64) a.dis[rowv + j] = INF32;
  if (lane == 0) a.dis[rowv + d - 1] = 0;
}

// Subtree tasks: a maximal subtree of at most a few hundred nodes depends only
// on itself and on ancestors finished by earlier launches, so one workgroup
// walks it level by level with workgroup barriers (its waves share one CU,
// hence one L1; no device-scope fence is needed between its levels).  The
// waves take (node, column pass) units of a level round-robin.
struct TaskArgs {
  LabelArgs a;
  const int32_t* lvl_base;   // [ntasks+1] first level of each task
  const int32_t* lvl_off;    // [levels+1] node offsets of every task level
  const int32_t* lvl_depth;  // [levels] depth of the level's nodes
  const int32_t* nodes;      // vertex ids, task by task, level by level
  int ntasks;
};

template <bool L4, bool IDX32>
__global__ void __launch_bounds__(1024) label_subtree_kernel(TaskArgs t) {
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = threadIdx.x & 63;
  const int chunk = (t.ntasks + XCDS - 1) / XCDS;
  const int task = (int)(blockIdx.x % XCDS) * chunk + (int)(blockIdx.x / XCDS);
  if (task >= t.ntasks) return;
  const int l1 = t.lvl_base[task + 1];
  for (int li = t.lvl_base[task]; li < l1; li++) {
    const int n0 = t.lvl_off[li], n1 = t.lvl_off[li + 1];
    const int d = t.lvl_depth[li];
    const int passes = d > 1 ? (d - 1 + 64 * R - 1) / (64 * R) : 1;
    const int units = (n1 - n0) * passes;
    for (int u = wid; u < units; u += nw) unit_row<L4, IDX32>(t.a, t.nodes[n0 + u / passes], d, (u % passes) * 64 * R, lane);
    __syncthreads();
  }
}

void launch_subtrees(const TaskArgs& t, hipStream_t s) {
  const dim3 grid((unsigned)(((t.ntasks + XCDS - 1) / XCDS) * XCDS)), block(1024);
  if (t.a.idx32) {
    if (t.a.l4) hipLaunchKernelGGL((label_subtree_kernel<true, true>), grid, block, 0, s, t);
    else hipLaunchKernelGGL((label_subtree_kernel<false, true>), grid, block, 0, s, t);
  } else {
    if (t.a.l4) hipLaunchKernelGGL((label_subtree_kernel<true, false>), grid, block, 0, s, t);
    else hipLaunchKernelGGL((label_subtree_kernel<false, false>), grid, block, 0, s, t);
  }
}

}  // namespace

#define H2H_TRY(x)               \
  do {                           \
    h2h_status s_ = (x);         \
    if (s_ != H2H_OK) return s_; \
  } while (0)

namespace {

// Work-item plan of one method.
//   PF       one item per node (its whole bag), 256 threads.
//   LL       one item per (node, bag entry), 64 threads, atomic combine.
//   BATCHED  per level: wide levels (>= 4 items per CU) one item per node;
//            narrow levels split large bags into groups so the level still
//            spreads over the chip; 256 threads.
h2h_status build_plan(h2h_ctx* ctx, int method, LabelPlan& P) {
  P.reset();
  const int levels = (int)ctx->level_off.size() - 1;
  const int n = ctx->n;
  // BATCHED: maximal subtrees of at most smax nodes become persistent tasks
  std::vector<int32_t> task_of(n, -1);
  if (method == H2H_METHOD_BATCHED) {
    const int64_t smax = std::max<int64_t>(64, n / (2LL * ctx->num_cus));
    std::vector<int64_t> sub(n, 1);
    for (int i = n - 1; i > 0; i--) {
      const int v = ctx->bfs[i];
      sub[ctx->father[v]] += sub[v];
    }
    int ntasks = 0;
    for (int i = 0; i < n; i++) {
      const int v = ctx->bfs[i];
      const int f = ctx->father[v];
      if (f >= 0 && task_of[f] >= 0) task_of[v] = task_of[f];
      else if (sub[v] <= smax && f >= 0) task_of[v] = ntasks++;
    }
    std::vector<int32_t> order;  // task nodes, task by task, BFS order inside
    std::vector<int32_t> cnt(ntasks + 1, 0);
    for (int i = 0; i < n; i++)
      if (task_of[ctx->bfs[i]] >= 0) cnt[task_of[ctx->bfs[i]] + 1]++;
    for (int t = 0; t < ntasks; t++) cnt[t + 1] += cnt[t];
    order.assign(cnt[ntasks], 0);
    {
      std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
      for (int i = 0; i < n; i++) {
        const int v = ctx->bfs[i];
        if (task_of[v] >= 0) order[fill[task_of[v]]++] = v;
      }
    }
    // levels of every task: start offsets into `order` and their depth
    std::vector<int32_t> lvl_base(1, 0), starts, lvl_depth;
    for (int t = 0; t < ntasks; t++) {
      for (int k = cnt[t]; k < cnt[t + 1]; k++) {
        if (k == cnt[t] || ctx->depth[order[k]] != ctx->depth[order[k - 1]]) {
          starts.push_back(k);
          lvl_depth.push_back(ctx->depth[order[k]]);
        }
      }
      lvl_base.push_back((int32_t)lvl_depth.size());
    }
    starts.push_back((int32_t)order.size());
    H2H_TRY(upload(ctx, P.task_lvl_base, lvl_base.data(), lvl_base.size() * sizeof(int32_t)));
    H2H_TRY(upload(ctx, P.task_lvl_off, starts.data(), starts.size() * sizeof(int32_t)));
    H2H_TRY(upload(ctx, P.task_lvl_depth, lvl_depth.data(), lvl_depth.size() * sizeof(int32_t)));
    H2H_TRY(upload(ctx, P.task_nodes, order.data(), order.size() * sizeof(int32_t)));
    P.ntasks = ntasks;
    P.task_nodes_n = (int64_t)order.size();
  }
  std::vector<Item> items;
  std::vector<int32_t> init_nodes, init_depth;
  P.lvl_off.assign(1, 0);
  const int64_t target = (int64_t)ctx->num_cus * 4;
  for (int l = 0; l < levels; l++) {
    const int b0 = ctx->level_off[l], b1 = ctx->level_off[l + 1];
    const int cnt = b1 - b0;
    int64_t entries = 0;
    int cnt_top = 0;
    for (int i = b0; i < b1; i++) {
      const int v = ctx->bfs[i];
      if (task_of[v] >= 0) continue;
      cnt_top++;
      entries += ctx->pos_hash[v + 1] - ctx->pos_hash[v] - 1;
    }
    int64_t group = INT32_MAX;  // entries per item
    int block = 256;
    if (method == H2H_METHOD_LL) {
      group = 1;
      block = 64;
    } else if (method == H2H_METHOD_BATCHED && cnt_top < target) {
      group = std::max<int64_t>(4, (entries + target - 1) / target);
    }
    (void)cnt;
    for (int i = b0; i < b1; i++) {
      const int v = ctx->bfs[i];
      if (task_of[v] >= 0) continue;
      const int32_t b = (int32_t)(ctx->pos_hash[v + 1] - ctx->pos_hash[v] - 1);
      const int32_t nitems = b == 0 ? 1 : (int32_t)((b + group - 1) / group);
      for (int32_t k = 0; k < nitems; k++) {
        Item it;
        it.node = i;
        it.first = (int32_t)(k * std::min<int64_t>(group, b));
        it.count = (int32_t)std::min<int64_t>(group, b - it.first);
        if (b == 0) it.first = it.count = 0;
        it.flags = (nitems == 1 ? ITEM_DIRECT : 0) | (k == 0 ? ITEM_LEAD : 0);
        items.push_back(it);
      }
      if (nitems > 1) {
        init_nodes.push_back(i);
        init_depth.push_back(l + 1);
      }
    }
    P.lvl_off.push_back((int32_t)items.size());
    P.lvl_block.push_back(block);
  }
  H2H_TRY(upload(ctx, P.items, items.data(), items.size() * sizeof(Item)));
  H2H_TRY(upload(ctx, P.init_nodes, init_nodes.data(), init_nodes.size() * sizeof(int32_t)));
  H2H_TRY(upload(ctx, P.init_depth, init_depth.data(), init_depth.size() * sizeof(int32_t)));
  P.n_items = (int64_t)items.size();
  P.n_init = (int64_t)init_nodes.size();
  P.method = method;
  return H2H_OK;
}

void enqueue_plan(const LabelPlan& P, const LabelArgs& a, hipStream_t s) {
  if (P.n_init > 0)
    hipLaunchKernelGGL(label_init_kernel, dim3((unsigned)((P.n_init + 3) / 4)), dim3(256), 0, s, a,
                       P.init_nodes.as<int32_t>(), P.init_depth.as<int32_t>(), (int)P.n_init);
  const int levels = (int)P.lvl_off.size() - 1;
  for (int l = 0; l < levels; l++) {
    const int cnt = P.lvl_off[l + 1] - P.lvl_off[l];
    if (cnt == 0) continue;
    launch_items(a, cnt, dim3(P.lvl_block[l]), s, P.items.as<Item>(), P.lvl_off[l], l + 1);
  }
  if (P.ntasks > 0) {
    TaskArgs t;
    t.a = a;
    t.lvl_base = P.task_lvl_base.as<int32_t>();
    t.lvl_off = P.task_lvl_off.as<int32_t>();
    t.lvl_depth = P.task_lvl_depth.as<int32_t>();
    t.nodes = P.task_nodes.as<int32_t>();
    t.ntasks = P.ntasks;
    launch_subtrees(t, s);
  }
}

}  // namespace

void invalidate_label_plans(h2h_ctx* ctx) {
  for (auto& P : ctx->lplan) P.reset();
}

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
  hipStream_t s = ctx->stream;
  const bool l4 = (kind == H2H_LABEL_L4);
  if (l4 && ctx->label_entries >= (int64_t)1 << 32)
    return set_err(ctx, H2H_ERR_OOM, "L4 build keeps 32-bit ancestor row offsets: index above 2^32 label entries");
  if (l4) H2H_TRY(dev_alloc(ctx, ctx->d_anc, std::max<int64_t>(ctx->label_entries, 1) * sizeof(uint32_t)));
  LabelArgs a;
  a.bfs = ctx->d_bfs.as<int32_t>();
  a.father = ctx->d_father.as<int32_t>();
  a.dis_hash = ctx->d_dis_hash.as<int64_t>();
  a.pos_hash = ctx->d_pos_hash.as<int64_t>();
  a.pos_pos = ctx->d_pos_pos.as<int32_t>();
  a.pos_w = ctx->d_pos_w.as<int32_t>();
  a.pos_rowoff = ctx->d_pos_rowoff.as<int64_t>();
  a.dis = ctx->d_dis.as<int32_t>();
  a.anc = l4 ? ctx->d_anc.as<uint32_t>() : nullptr;
  a.l4 = l4 ? 1 : 0;
  a.idx32 = ctx->label_entries < ((int64_t)1 << 30) ? 1 : 0;
  LabelPlan& P = ctx->lplan[method];
  if (P.method != method) H2H_TRY(build_plan(ctx, method, P));
  if (method == H2H_METHOD_BATCHED && !P.graph[l4]) {
    // capture the whole level sweep once; replays skip per-launch host work
    hipGraph_t g = nullptr;
    H2H_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_plan(P, a, s);
    H2H_HIP(ctx, hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(&P.graph[l4], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_err(ctx, e, "hipGraphInstantiate");
  }
  hipEvent_t e0, e1;
  H2H_HIP(ctx, hipEventCreate(&e0));
  H2H_HIP(ctx, hipEventCreate(&e1));
  H2H_HIP(ctx, hipEventRecord(e0, s));
  if (method == H2H_METHOD_BATCHED) {
    H2H_HIP(ctx, hipGraphLaunch(P.graph[l4], s));
  } else {
    enqueue_plan(P, a, s);
  }
  H2H_HIP(ctx, hipGetLastError());
  H2H_HIP(ctx, hipEventRecord(e1, s));
  H2H_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0;
  H2H_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->t_label_us = ms * 1000.0;
  if (usec) *usec = ctx->t_label_us;
  ctx->label_kind = kind;
  ctx->labels_ready = true;
  return H2H_OK;
}

h2h_status profile_label_build(h2h_ctx* ctx, int kind, int method, double* kernel_us, int64_t* launches) {
  hipStream_t s = ctx->stream;
  const bool l4 = (kind == H2H_LABEL_L4);
  if (l4 && ctx->label_entries >= (int64_t)1 << 32)
    return set_err(ctx, H2H_ERR_OOM, "L4 build keeps 32-bit ancestor row offsets: index above 2^32 label entries");
  if (l4) H2H_TRY(dev_alloc(ctx, ctx->d_anc, std::max<int64_t>(ctx->label_entries, 1) * sizeof(uint32_t)));
  LabelArgs a;
  a.bfs = ctx->d_bfs.as<int32_t>();
  a.father = ctx->d_father.as<int32_t>();
  a.dis_hash = ctx->d_dis_hash.as<int64_t>();
  a.pos_hash = ctx->d_pos_hash.as<int64_t>();
  a.pos_pos = ctx->d_pos_pos.as<int32_t>();
  a.pos_w = ctx->d_pos_w.as<int32_t>();
  a.pos_rowoff = ctx->d_pos_rowoff.as<int64_t>();
  a.dis = ctx->d_dis.as<int32_t>();
  a.anc = l4 ? ctx->d_anc.as<uint32_t>() : nullptr;
  a.l4 = l4 ? 1 : 0;
  a.idx32 = ctx->label_entries < ((int64_t)1 << 30) ? 1 : 0;
  LabelPlan& P = ctx->lplan[method];
  if (P.method != method) H2H_TRY(build_plan(ctx, method, P));
  const int levels = (int)P.lvl_off.size() - 1;
  std::vector<hipEvent_t> ev(2 * levels + 2, nullptr);
  for (auto& e : ev) H2H_HIP(ctx, hipEventCreate(&e));
  int64_t nl = 0;
  if (P.n_init > 0) {
    H2H_HIP(ctx, hipEventRecord(ev[2 * levels], s));
    hipLaunchKernelGGL(label_init_kernel, dim3((unsigned)((P.n_init + 3) / 4)), dim3(256), 0, s, a,
                       P.init_nodes.as<int32_t>(), P.init_depth.as<int32_t>(), (int)P.n_init);
    H2H_HIP(ctx, hipEventRecord(ev[2 * levels + 1], s));
  }
  for (int l = 0; l < levels; l++) {
    const int cnt = P.lvl_off[l + 1] - P.lvl_off[l];
    H2H_HIP(ctx, hipEventRecord(ev[2 * l], s));
    if (cnt > 0) {
      launch_items(a, cnt, dim3(P.lvl_block[l]), s, P.items.as<Item>(), P.lvl_off[l], l + 1);
      nl++;
    }
    H2H_HIP(ctx, hipEventRecord(ev[2 * l + 1], s));
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (P.ntasks > 0) {
    H2H_HIP(ctx, hipEventCreate(&t0));
    H2H_HIP(ctx, hipEventCreate(&t1));
    TaskArgs t;
    t.a = a;
    t.lvl_base = P.task_lvl_base.as<int32_t>();
    t.lvl_off = P.task_lvl_off.as<int32_t>();
    t.lvl_depth = P.task_lvl_depth.as<int32_t>();
    t.nodes = P.task_nodes.as<int32_t>();
    t.ntasks = P.ntasks;
    H2H_HIP(ctx, hipEventRecord(t0, s));
    launch_subtrees(t, s);
    H2H_HIP(ctx, hipEventRecord(t1, s));
    nl++;
  }
  H2H_HIP(ctx, hipGetLastError());
  H2H_HIP(ctx, hipStreamSynchronize(s));
  double tot = 0;
  for (int l = 0; l < levels; l++) {
    float ms = 0;
    H2H_HIP(ctx, hipEventElapsedTime(&ms, ev[2 * l], ev[2 * l + 1]));
    tot += ms * 1000.0;
  }
  if (t0) {
    float ms = 0;
    H2H_HIP(ctx, hipEventElapsedTime(&ms, t0, t1));
    tot += ms * 1000.0;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (kernel_us) *kernel_us = tot;
  if (launches) *launches = nl;
  ctx->label_kind = kind;
  ctx->labels_ready = true;
  return H2H_OK;
}

}  // namespace h2h
