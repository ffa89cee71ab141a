// Spaxel assignment and the stable sort-by-spaxel index.
//
// rbx_spaxel_assign replaces square_spaxel_assignment
// (rubix/telescope/utils.py:77-152).  rbx_spaxel_sort is new: a stable,
// deterministic counting sort of the spaxel ids (the reference analogue is
// the lexsort + searchsorted of rubix/spectra/dust/dust_extinction.py:237-267).
//
// Sort design (integer work, HBM-bound, no floating point):
//   keys are bucketed into B = n_keys + 1 buckets (the last one collects ids
//   outside [0, n_keys)); the input is cut into tiles, each tile owned by one
//   wavefront that walks it in order, 64 stars per round.
//   1. count  : per (bucket, tile) counts.  Lanes with equal keys in a round
//               are found with one ballot per key bit; the lowest lane adds
//               the group size.  Only the tile's own wavefront touches a
//               (bucket, tile) counter, so the result is deterministic.
//   2. scan   : exclusive scan over the bucket-major (bucket, tile) table
//               gives each tile its base position inside each bucket.
//   3. scatter: the same walk; position = base + rank inside the round's
//               equal-key group, so stars keep their input order (stable).
//   The (bucket, tile) table of 4096-star tiles is kept to 2^24 entries.  Key
//   counts past that (a batch of many galaxies: 310 000 keys over 31.8 M
//   stars) sort in LSD radix passes over D-bit digits of the bucket index,
//   each pass the same stable count + scan + scatter over 2^D buckets, the
//   scatter carrying each star's input index and bucket along; stable
//   passes give the same order as the one-pass sort.  Bucket offsets then
//   come from a binary search of the sorted buckets.
#include <stdarg.h>
#include <string.h>

#include "rbx_common.h"

namespace rbx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return RBX_EHIP;
  }
  return RBX_OK;
}

// ------------------------------------------------------------ spaxel assign

template <bool kLds>
__global__ void spaxel_assign_kernel(const float* __restrict__ coords, int64_t n,
                                     int64_t stride, const float* __restrict__ edges,
                                     int n_edges, int32_t* __restrict__ out) {
  extern __shared__ float s_edges[];
  const float* e = edges;
  if (kLds) {
    for (int k = threadIdx.x; k < n_edges; k += blockDim.x) s_edges[k] = edges[k];
    __syncthreads();
    e = s_edges;
  }
  const int nb = n_edges - 1;
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float x = coords[i * stride];
    const float y = coords[i * stride + 1];
    int ix = upper_bound_f(e, n_edges, x) - 1;
    int iy = upper_bound_f(e, n_edges, y) - 1;
    ix = ix < 0 ? 0 : (ix > nb - 1 ? nb - 1 : ix);
    iy = iy < 0 ? 0 : (iy > nb - 1 ? nb - 1 : iy);
    out[i] = ix + nb * iy;
  }
}

// ------------------------------------------------------------ integer scan

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanBlock = kScanThreads * kScanItems;

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = gtid(); i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}

template <class T>
__device__ inline T block_exclusive_scan(T v, T* smem, T* total) {
  // smem: kScanThreads entries.  Hillis-Steele over the block (deterministic).
  const int t = threadIdx.x;
  smem[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    T add = (t >= off) ? smem[t - off] : T(0);
    __syncthreads();
    smem[t] += add;
    __syncthreads();
  }
  T incl = smem[t];
  *total = smem[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ void scan_reduce_kernel(const int32_t* __restrict__ in, int64_t n,
                                   int32_t* __restrict__ block_sums) {
  __shared__ int32_t smem[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanBlock;
  int32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + int64_t(k) * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  int32_t tot;
  block_exclusive_scan<int32_t>(s, smem, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the block sums in place.
__global__ void scan_top_kernel(int32_t* __restrict__ sums, int64_t nblk) {
  __shared__ int32_t smem[kScanThreads];
  int32_t carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += kScanThreads) {
    int64_t i = b0 + threadIdx.x;
    int32_t v = i < nblk ? sums[i] : 0;
    int32_t tot;
    int32_t ex = block_exclusive_scan<int32_t>(v, smem, &tot);
    if (i < nblk) sums[i] = carry + ex;
    carry += tot;
  }
}

// Each block: exclusive scan of its kScanBlock elements (thread-contiguous
// runs of kScanItems) plus the block offset, in place.
__global__ void scan_down_kernel(int32_t* __restrict__ data, int64_t n,
                                 const int32_t* __restrict__ block_off) {
  __shared__ int32_t smem[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? data[i] : 0;
    s += v[k];
  }
  int32_t tot;
  int32_t ex = block_exclusive_scan<int32_t>(s, smem, &tot) + block_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) data[i] = ex;
    ex += v[k];
  }
}

// ------------------------------------------------------------ sort

struct SortPlan {
  int64_t n;
  int32_t n_keys;
  int32_t buckets;  // n_keys + 1 (overflow bucket last)
  int64_t tile;     // stars per tile, multiple of 64
  int64_t ntiles;
  int64_t table;    // (buckets or digit buckets) * ntiles
  int64_t nblk;     // scan blocks over the table
  int nbits;        // bits needed to tell buckets apart
  int digit_bits;   // 0: one pass over all buckets; else radix digits of this width
  int passes;       // counting-sort passes
};

constexpr int64_t kTableCap = int64_t(1) << 24;  // (bucket, tile) counters
constexpr int64_t kSortTile = 4096;              // stars per tile

static SortPlan make_plan(int64_t n, int32_t n_keys) {
  SortPlan p;
  p.n = n;
  p.n_keys = n_keys;
  p.buckets = n_keys + 1;
  p.tile = kSortTile;
  p.ntiles = n > 0 ? (n + p.tile - 1) / p.tile : 1;
  int b = 1;
  while ((int64_t(1) << b) < p.buckets) ++b;
  p.nbits = b;
  if (int64_t(p.buckets) * p.ntiles <= kTableCap) {
    p.digit_bits = 0;
    p.passes = 1;
    p.table = int64_t(p.buckets) * p.ntiles;
  } else {
    // ntiles <= 2^19 for n < 2^31, so digits of at least 5 bits fit
    int d = 16;
    while (d > 1 && (int64_t(1) << d) * p.ntiles > kTableCap) --d;
    p.digit_bits = d;
    p.passes = (p.nbits + d - 1) / d;
    p.table = (int64_t(1) << d) * p.ntiles;
  }
  p.nblk = (p.table + kScanBlock - 1) / kScanBlock;
  return p;
}

// Radix passes keep the input index and bucket of every position: three
// n-element buffers (one index, two buckets) beside the caller's perm.
template <class S>
static void sort_layout(S& c, const SortPlan& p, int32_t** table, int32_t** blocks,
                        int32_t** tperm, int32_t** b0, int32_t** b1) {
  auto* t = c.template take<int32_t>(p.table);
  auto* bl = c.template take<int32_t>(p.nblk + 1);
  const int64_t rn = p.digit_bits ? (p.n > 0 ? p.n : 1) : 0;
  auto* tp = c.template take<int32_t>(rn);
  auto* x0 = c.template take<int32_t>(rn);
  auto* x1 = c.template take<int32_t>(rn);
  if (table) *table = reinterpret_cast<int32_t*>(t);
  if (blocks) *blocks = reinterpret_cast<int32_t*>(bl);
  if (tperm) *tperm = reinterpret_cast<int32_t*>(tp);
  if (b0) *b0 = reinterpret_cast<int32_t*>(x0);
  if (b1) *b1 = reinterpret_cast<int32_t*>(x1);
}

struct SizerAdapter {
  Sizer s;
  template <class T>
  T* take(size_t count) {
    s.take<T>(count);
    return nullptr;
  }
};

__device__ inline int bucket_of(int32_t id, int32_t n_keys) {
  return (id >= 0 && id < n_keys) ? id : n_keys;
}

// Lanes of the wavefront holding the same key (restricted to `active`).
__device__ inline uint64_t peer_mask(int key, uint64_t active, int nbits) {
  uint64_t peers = active;
  for (int b = 0; b < nbits; ++b) {
    const int bit = (key >> b) & 1;
    const uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

// One counting-sort pass.  The sort key of element i is its bucket
// bucket_of(in[i]), or with mask != 0 the digit (bucket >> shift) & mask.
// Scatter: perm[pos] = the element's input index (src_perm[i], or i on the
// first pass), bucket_out[pos] = its bucket, sorted[pos] = its original id
// (ids[perm[pos]]; in[i] itself on the first pass).
template <bool kScatter>
__global__ void __launch_bounds__(64) sort_walk_kernel(
    const int32_t* __restrict__ in, int64_t n, int32_t n_keys, int64_t tile,
    int64_t ntiles, int nbits, int shift, int32_t mask, int32_t* __restrict__ table,
    const int32_t* __restrict__ src_perm, const int32_t* __restrict__ ids,
    int32_t* __restrict__ perm, int32_t* __restrict__ bucket_out,
    int32_t* __restrict__ sorted) {
  // blockDim.x == 64: one wavefront walks one tile in input order.
  const int lane = threadIdx.x;
  const int64_t t = blockIdx.x;
  const int64_t begin = t * tile;
  const int64_t end = begin + tile < n ? begin + tile : n;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int64_t r = begin; r < end; r += 64) {
    const int64_t i = r + lane;
    const bool valid = i < end;
    const int32_t id = valid ? in[i] : 0;
    const int full = bucket_of(id, n_keys);
    const int key = mask ? ((full >> shift) & mask) : full;
    const uint64_t active = __ballot(valid);
    const uint64_t peers = peer_mask(key, active, nbits);  // all lanes take part
    const int leader = __ffsll((unsigned long long)peers) - 1;
    const int cnt = __popcll(peers);
    const int rank = __popcll(peers & lt_mask);
    int32_t* slot = table + int64_t(key) * ntiles + t;
    int32_t base = 0;
    if (valid && lane == leader) base = atomicAdd(slot, cnt);
    if (kScatter) {
      base = __shfl(base, leader < 0 ? 0 : leader, 64);
      if (valid) {
        const int32_t pos = base + rank;
        const int32_t src = src_perm ? src_perm[i] : int32_t(i);
        perm[pos] = src;
        if (bucket_out) bucket_out[pos] = full;
        if (sorted) sorted[pos] = src_perm ? ids[src] : id;
      }
    }
  }
}

// offsets[k] = number of sorted buckets below k, k in [0, n_keys].
__global__ void bucket_offsets_kernel(const int32_t* __restrict__ bucket, int64_t n,
                                      int32_t n_keys, int32_t* __restrict__ offsets) {
  for (int64_t k = gtid(); k <= n_keys; k += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (bucket[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    offsets[k] = int32_t(lo);
  }
}

__global__ void sort_offsets_kernel(const int32_t* __restrict__ table, int32_t n_keys,
                                    int64_t ntiles, int32_t* __restrict__ offsets) {
  for (int64_t k = gtid(); k <= n_keys; k += int64_t(gridDim.x) * blockDim.x)
    offsets[k] = table[k * ntiles];
}

static int grid_for(int64_t n, int threads, int64_t cap = 65536) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

int exclusive_scan_i32(int32_t* data, int64_t n, int32_t* blocks, hipStream_t s) {
  const int64_t nblk = (n + kScanBlock - 1) / kScanBlock;
  if (nblk == 0) return RBX_OK;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(nblk), dim3(kScanThreads), 0, s, data, n, blocks);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanThreads), 0, s, blocks, nblk);
  hipLaunchKernelGGL(scan_down_kernel, dim3(nblk), dim3(kScanThreads), 0, s, data, n, blocks);
  return check_launch("exclusive_scan_i32");
}

void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}

size_t sort_workspace_size(int64_t n, int32_t n_keys) {
  SortPlan p = make_plan(n, n_keys);
  SizerAdapter a;
  sort_layout(a, p, nullptr, nullptr, nullptr, nullptr, nullptr);
  return a.s.used;
}

// Workspace enough for any key count in [1, max_keys]: the table never
// passes 2^24 entries, and the radix buffers are there whenever max_keys
// would need them.
size_t sort_workspace_bound(int64_t n, int32_t max_keys) {
  SortPlan p = make_plan(n, max_keys);
  p.table = kTableCap;
  p.nblk = (p.table + kScanBlock - 1) / kScanBlock;
  SizerAdapter a;
  sort_layout(a, p, nullptr, nullptr, nullptr, nullptr, nullptr);
  return a.s.used;
}

// Count + scan of one pass; returns the launch status.
static int sort_pass(const SortPlan& p, const int32_t* in, int shift, int32_t mask, int nbits,
                     int32_t* table, int32_t* blocks, int64_t table_n, const int32_t* src_perm,
                     const int32_t* ids, int32_t* perm, int32_t* bucket_out, int32_t* sorted,
                     int32_t* offsets, hipStream_t s) {
  fill_i32(table, table_n, 0, s);
  if (p.n > 0) {
    hipLaunchKernelGGL(sort_walk_kernel<false>, dim3(p.ntiles), dim3(64), 0, s, in, p.n,
                       p.n_keys, p.tile, p.ntiles, nbits, shift, mask, table, nullptr, nullptr,
                       nullptr, nullptr, nullptr);
  }
  int rc = exclusive_scan_i32(table, table_n, blocks, s);
  if (rc) return rc;
  if (offsets)
    hipLaunchKernelGGL(sort_offsets_kernel, dim3(grid_for(int64_t(p.n_keys) + 1, 256)), dim3(256),
                       0, s, table, p.n_keys, p.ntiles, offsets);
  if (p.n > 0) {
    hipLaunchKernelGGL(sort_walk_kernel<true>, dim3(p.ntiles), dim3(64), 0, s, in, p.n, p.n_keys,
                       p.tile, p.ntiles, nbits, shift, mask, table, src_perm, ids, perm,
                       bucket_out, sorted);
  }
  return check_launch("rbx_spaxel_sort");
}

// Sort used by the fused path too; the workspace is carved by the caller.
int spaxel_sort_impl(const int32_t* ids, int64_t n, int32_t n_keys, int32_t* perm,
                     int32_t* sorted, int32_t* offsets, void* ws, size_t ws_bytes,
                     hipStream_t s) {
  SortPlan p = make_plan(n, n_keys);
  Carver c(ws, ws_bytes);
  int32_t *table = nullptr, *blocks = nullptr, *tperm = nullptr, *b0 = nullptr, *b1 = nullptr;
  sort_layout(c, p, &table, &blocks, &tperm, &b0, &b1);
  if (!c.ok) return fail(RBX_EWORKSPACE, "rbx_spaxel_sort: workspace too small");
  if (!p.digit_bits)
    return sort_pass(p, ids, 0, 0, p.nbits, table, blocks, p.table, nullptr, ids, perm, nullptr,
                     sorted, offsets, s);
  // LSD radix: pass q sorts by bits [q D, q D + D) of the bucket, stably;
  // the index buffers alternate so that the last pass writes perm
  const int32_t mask = (int32_t(1) << p.digit_bits) - 1;
  const int32_t* in = ids;
  const int32_t* src = nullptr;
  int32_t* bufs[2] = {b0, b1};
  for (int q = 0; q < p.passes; ++q) {
    const bool last = q == p.passes - 1;
    int32_t* pout = ((p.passes - 1 - q) & 1) ? tperm : perm;
    int32_t* bout = bufs[q & 1];
    const int width = last ? p.nbits - q * p.digit_bits : p.digit_bits;
    int rc = sort_pass(p, in, q * p.digit_bits, mask, width, table, blocks, p.table, src, ids,
                       pout, bout, last ? sorted : nullptr, nullptr, s);
    if (rc) return rc;
    in = bout;
    src = pout;
  }
  hipLaunchKernelGGL(bucket_offsets_kernel, dim3(grid_for(int64_t(n_keys) + 1, 256)), dim3(256), 0,
                     s, in, n, n_keys, offsets);
  return check_launch("rbx_spaxel_sort");
}

}  // namespace rbx

extern "C" {

int rbx_version(void) { return 100; }

const char* rbx_last_error_string(void) { return rbx::g_err; }

int rbx_spaxel_assign(const float* coords, int64_t n, int64_t stride, const float* edges,
                      int32_t n_edges, int32_t* out_ids, void* stream) {
  using namespace rbx;
  if (n < 0 || stride < 2 || n_edges < 2 || (n > 0 && (!coords || !out_ids)) || !edges)
    return fail(RBX_EINVAL, "rbx_spaxel_assign: bad arguments (need n >= 0, stride >= 2, n_edges >= 2)");
  if (n == 0) return RBX_OK;
  hipStream_t s = as_stream(stream);
  const int threads = 256;
  const int grid = grid_for(n, threads);
  if (n_edges <= 8192) {
    hipLaunchKernelGGL(spaxel_assign_kernel<true>, dim3(grid), dim3(threads),
                       size_t(n_edges) * sizeof(float), s, coords, n, stride, edges, n_edges, out_ids);
  } else {
    hipLaunchKernelGGL(spaxel_assign_kernel<false>, dim3(grid), dim3(threads), 0, s, coords, n,
                       stride, edges, n_edges, out_ids);
  }
  return check_launch("rbx_spaxel_assign");
}

size_t rbx_spaxel_sort_workspace_size(int64_t n, int32_t n_keys) {
  if (n < 0 || n_keys < 1) return 0;
  return rbx::sort_workspace_size(n, n_keys);
}

int rbx_spaxel_sort(const int32_t* ids, int64_t n, int32_t n_keys, int32_t* perm,
                    int32_t* sorted, int32_t* offsets, void* workspace, size_t workspace_bytes,
                    void* stream) {
  using namespace rbx;
  if (n < 0 || n_keys < 1 || !offsets || (n > 0 && (!ids || !perm)))
    return fail(RBX_EINVAL, "rbx_spaxel_sort: bad arguments");
  if (n > INT32_MAX - 1) return fail(RBX_EINVAL, "rbx_spaxel_sort: n must fit int32");
  if (int64_t(n_keys) + 1 > (int64_t(1) << 30))
    return fail(RBX_EINVAL, "rbx_spaxel_sort: n_keys too large");
  return spaxel_sort_impl(ids, n, n_keys, perm, sorted, offsets, workspace, workspace_bytes,
                          as_stream(stream));
}

}  // extern "C"
