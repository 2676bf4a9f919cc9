// bs_conn.hip -- connectivity sampled on the device (gfx950): the random rules
// fixed_indegree (with multapses) and pairwise_bernoulli, written straight into
// device memory in the reference's edge order, so a network's edges never
// cross PCIe before bs_net_create builds its gather segments from them.
//
// Reference samplers: sample_fixed_indegree / sample_pairwise_bernoulli
// (_nest_network/connectivity.py:81-140). The reference draws with
// jax.random (fold_in per post + choice; uniform over the pair grid); here
// every draw is one Threefry-2x32-20 block keyed by the projection seed, so
// each edge is a pure function of (seed, post, slot) or (seed, pre, post):
//   key     = (seed & 0xFFFFFFFF, 0xE36F6E6E ^ (seed >> 32))
//   fixed in-degree, post j, slot k < K: (x0, x1) = threefry(key, (j, k)),
//     r = x1 * 2^32 + x0, c = floor(r * n_cand / 2^64); with self-exclusion
//     n_cand = n_pre - 1 and pre = c + (c >= j) (the candidates
//     arange(0, j) ++ arange(j + 1, n_pre) of connectivity.py:117-121),
//     else n_cand = n_pre and pre = c;
//   pairwise Bernoulli, pair (i, j): x0 of threefry(key, (i, j)); the pair is
//     connected iff x0 < ceil(p * 2^32), i.e. the uniform x0 * 2^-32 < p
//     (connectivity.py:92), with i == j skipped under self-exclusion; edges
//     in row-major (pre, then post) order, as jnp.where returns them.
// Optional maps translate segment-local indices to population-local ones.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_reduce.hpp>
#include <hipcub/device/device_scan.hpp>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../../include/brainspike_hip.h"
#include "bs_math.h"

int bs_internal_fail(int code, const std::string& msg);

// Scratch of bs_conn_bernoulli_count_noalloc (16 MiB per GPU, see there).
constexpr size_t kScratchBytes = 16u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using bs::threefry2x32;

#define CCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return bs_internal_fail(BS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr uint32_t kConnTag = 0xE36F6E6Eu;

__global__ void k_conn_indegree(uint32_t k0, uint32_t k1, long long E, int K, uint32_t n_cand, int excl,
                                const int32_t* __restrict__ pre_map, const int32_t* __restrict__ post_map,
                                int32_t* __restrict__ pre, int32_t* __restrict__ post) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const uint32_t j = (uint32_t)(e / K), k = (uint32_t)(e - (long long)j * K);
  uint32_t x0, x1;
  threefry2x32(k0, k1, j, k, x0, x1);
  const uint64_t r = ((uint64_t)x1 << 32) | x0;
  uint32_t c = (uint32_t)__umul64hi(r, (uint64_t)n_cand);
  if (excl) c += (c >= j) ? 1u : 0u;
  pre[e] = pre_map ? pre_map[c] : (int32_t)c;
  post[e] = post_map ? post_map[j] : (int32_t)j;
}

// One block per pre row i: connected pairs of the row, and per-post counts.
__global__ void k_conn_bern_count(uint32_t k0, uint32_t k1, int n_post, uint64_t thr, int excl,
                                  long long* __restrict__ row_cnt, int* __restrict__ col_cnt) {
  const uint32_t i = blockIdx.x;
  __shared__ int part[4];
  int n = 0;
  for (int j = threadIdx.x; j < n_post; j += 256) {
    uint32_t x0, x1;
    threefry2x32(k0, k1, i, (uint32_t)j, x0, x1);
    const bool on = (uint64_t)x0 < thr && !(excl && (uint32_t)j == i);
    if (on) {
      ++n;
      atomicAdd(&col_cnt[j], 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) row_cnt[i] = (long long)part[0] + part[1] + part[2] + part[3];
}

// One block per pre row i: the row's connected posts in ascending order,
// written from row_ptr[i] (exclusive prefix of the count pass).
__global__ void k_conn_bern_fill(uint32_t k0, uint32_t k1, int n_post, uint64_t thr, int excl,
                                 const long long* __restrict__ row_ptr, const int32_t* __restrict__ pre_map,
                                 const int32_t* __restrict__ post_map, int32_t* __restrict__ pre,
                                 int32_t* __restrict__ post) {
  const uint32_t i = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __shared__ int wcnt[4];
  long long out = row_ptr[i];
  const int32_t pre_v = pre_map ? pre_map[i] : (int32_t)i;
  for (int j0 = 0; j0 < n_post; j0 += 256) {
    const int j = j0 + tid;
    bool on = false;
    if (j < n_post) {
      uint32_t x0, x1;
      threefry2x32(k0, k1, i, (uint32_t)j, x0, x1);
      on = (uint64_t)x0 < thr && !(excl && (uint32_t)j == i);
    }
    const unsigned long long m = __ballot(on);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wcnt[w];
    const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    if (on) {
      pre[out + base + below] = pre_v;
      post[out + base + below] = post_map ? post_map[j] : (int32_t)j;
    }
    out += total;
    __syncthreads();  // wcnt is rewritten by the next chunk
  }
}

void conn_key(int64_t seed, uint32_t& k0, uint32_t& k1) {
  k0 = (uint32_t)((uint64_t)seed & 0xFFFFFFFFull);
  k1 = kConnTag ^ (uint32_t)((uint64_t)seed >> 32);
}

// ceil(p * 2^32) clamped to [0, 2^32]: x0 < thr  <=>  x0 * 2^-32 < p.
uint64_t bern_threshold(double p) {
  if (!(p > 0.0)) return 0;
  if (p >= 1.0) return 1ull << 32;
  return (uint64_t)ceil(p * 4294967296.0);
}

// ---- scratch of bs_conn_bernoulli_count_noalloc: a file-scope device array
// (each GPU holds its own copy), addressed from the host only through the
// pointer hipGetSymbolAddress returns for the current device. Layout for a
// call: row counts int64 [n_pre + 1], chunk sums int64 [ceil((n_pre + 1) /
// 1024)], post counts int32 [n_post], largest post count int32 [1], each part
// 256-byte aligned.
constexpr int kNoAllocMaxRows = 1 << 20;  // n_pre and n_post limits
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

__global__ void k_fill_i32(int32_t* __restrict__ p, long long n, int32_t v) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}
__global__ void k_fill_i64(long long* __restrict__ p, long long n, long long v) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}

// Exclusive prefix sum over n int64 values, chunks of 1024 per block (four
// per thread): out[k] = in[0] + ... + in[k - 1] within the chunk, and the
// chunk's total into csum[chunk].
__global__ __launch_bounds__(256) void k_scan_chunks(const long long* __restrict__ in, long long n,
                                                     long long* __restrict__ out, long long* __restrict__ csum) {
  __shared__ long long wsum[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long base = (long long)blockIdx.x * 1024 + (long long)tid * 4;
  long long v[4], t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    t += v[k];
  }
  long long inc = t;  // inclusive scan of the thread totals over the wave
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  long long woff = 0;
  for (int w = 0; w < wave; ++w) woff += wsum[w];
  long long ex = woff + inc - t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (tid == 255) csum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// In place exclusive prefix of the nc chunk totals: one block, 256 at a time
// with a running carry.
__global__ __launch_bounds__(256) void k_scan_csum(long long* __restrict__ csum, int nc) {
  __shared__ long long wsum[4];
  __shared__ long long carry;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nc; c0 += 256) {
    const long long t = c0 + tid < nc ? csum[c0 + tid] : 0;
    long long inc = t;
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    long long woff = carry;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    if (c0 + tid < nc) csum[c0 + tid] = woff + inc - t;
    __syncthreads();
    if (tid == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_scan_add(long long* __restrict__ out, long long n,
                                                  const long long* __restrict__ csum) {
  const long long base = (long long)blockIdx.x * 1024 + (long long)threadIdx.x * 4;
  const long long add = csum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) out[base + k] += add;
}

// Largest of n non-negative int32 values into *out (which starts at 0).
__global__ __launch_bounds__(256) void k_max_i32(const int32_t* __restrict__ x, long long n, int32_t* __restrict__ out) {
  int m = 0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    m = max(m, x[k]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace

extern "C" {

int bs_conn_fixed_indegree(int64_t seed, int32_t n_pre, int32_t n_post, int32_t K, int32_t exclude_self,
                           const int32_t* pre_map, const int32_t* post_map, int32_t* pre, int32_t* post) {
  if (n_pre < 0 || n_post < 0 || K < 0) return bs_internal_fail(BS_EINVAL, "bs_conn_fixed_indegree: negative size");
  const long long E = (long long)n_post * K;
  if (E == 0) return BS_OK;
  if (E >= (1ll << 32)) return bs_internal_fail(BS_EINVAL, "bs_conn_fixed_indegree: more than 2^32 - 1 edges");
  if (!pre || !post) return bs_internal_fail(BS_EINVAL, "bs_conn_fixed_indegree: null output");
  const long long n_cand = exclude_self ? (long long)n_pre - 1 : (long long)n_pre;
  if (n_cand <= 0) return bs_internal_fail(BS_EINVAL, "bs_conn_fixed_indegree: no candidate presynaptic neurons");
  uint32_t k0, k1;
  conn_key(seed, k0, k1);
  hipLaunchKernelGGL(k_conn_indegree, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, 0, k0, k1, E, (int)K,
                     (uint32_t)n_cand, exclude_self ? 1 : 0, pre_map, post_map, pre, post);
  CCHK(hipGetLastError());
  CCHK(hipStreamSynchronize(0));
  return BS_OK;
}

int bs_conn_bernoulli_count(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                            int64_t* row_ptr, int64_t* n_edges, int32_t* max_indegree) {
  if (n_pre < 0 || n_post < 0 || !(p >= 0.0)) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count: bad sizes or p");
  if (!row_ptr || !n_edges || !max_indegree) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count: null argument");
  *n_edges = 0;
  *max_indegree = 0;
  if (n_pre == 0) return hipMemset(row_ptr, 0, sizeof(int64_t)) == hipSuccess ? BS_OK
                                                                               : bs_internal_fail(BS_EHIP, "hipMemset");
  uint32_t k0, k1;
  conn_key(seed, k0, k1);
  const uint64_t thr = bern_threshold(p);
  long long* cnt = nullptr;
  int* col = nullptr;
  int* colmax = nullptr;
  void* tmp = nullptr;
  size_t tb_scan = 0, tb_max = 0;
  hipError_t e = hipMalloc(&cnt, sizeof(long long) * ((size_t)n_pre + 1));
  if (e == hipSuccess) e = hipMalloc(&col, sizeof(int) * std::max<size_t>(1, (size_t)n_post));
  if (e == hipSuccess) e = hipMalloc(&colmax, sizeof(int));
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, cnt, (long long*)row_ptr, n_pre + 1);
  if (e == hipSuccess) e = hipcub::DeviceReduce::Max(nullptr, tb_max, col, colmax, std::max(1, n_post));
  if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(256, std::max(tb_scan, tb_max)));
  if (e == hipSuccess) e = hipMemset(cnt, 0, sizeof(long long) * ((size_t)n_pre + 1));
  if (e == hipSuccess) e = hipMemset(col, 0, sizeof(int) * std::max<size_t>(1, (size_t)n_post));
  if (e == hipSuccess && n_post > 0) {
    hipLaunchKernelGGL(k_conn_bern_count, dim3((unsigned)n_pre), dim3(256), 0, 0, k0, k1, (int)n_post, thr,
                       exclude_self ? 1 : 0, cnt, col);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tb_scan, cnt, (long long*)row_ptr, n_pre + 1);
  if (e == hipSuccess) e = hipcub::DeviceReduce::Max(tmp, tb_max, col, colmax, std::max(1, n_post));
  long long total = 0;
  int mx = 0;
  if (e == hipSuccess) e = hipMemcpy(&total, row_ptr + n_pre, sizeof(long long), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&mx, colmax, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(cnt);
  (void)hipFree(col);
  (void)hipFree(colmax);
  (void)hipFree(tmp);
  if (e != hipSuccess) return bs_internal_fail(BS_EHIP, std::string("bs_conn_bernoulli_count: ") + hipGetErrorString(e));
  *n_edges = total;
  *max_indegree = mx;
  return BS_OK;
}

int bs_conn_bernoulli_fill(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                           const int64_t* row_ptr, const int32_t* pre_map, const int32_t* post_map,
                           int32_t* pre, int32_t* post) {
  if (n_pre < 0 || n_post < 0 || !(p >= 0.0)) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_fill: bad sizes or p");
  if (!row_ptr) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_fill: null row_ptr");
  if (n_pre == 0 || n_post == 0 || !(p > 0.0)) return BS_OK;
  if (!pre || !post) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_fill: null output");
  uint32_t k0, k1;
  conn_key(seed, k0, k1);
  hipLaunchKernelGGL(k_conn_bern_fill, dim3((unsigned)n_pre), dim3(256), 0, 0, k0, k1, (int)n_post,
                     bern_threshold(p), exclude_self ? 1 : 0, (const long long*)row_ptr, pre_map, post_map, pre,
                     post);
  CCHK(hipGetLastError());
  CCHK(hipStreamSynchronize(0));
  return BS_OK;
}

int bs_conn_bernoulli_count_noalloc(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                                    int64_t* row_ptr, int64_t* n_edges, int32_t* max_indegree) {
  if (n_pre < 0 || n_post < 0 || !(p >= 0.0))
    return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count_noalloc: bad sizes or p");
  if (n_pre > kNoAllocMaxRows || n_post > kNoAllocMaxRows)
    return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count_noalloc: n_pre and n_post must be at most " +
                                           std::to_string(kNoAllocMaxRows) + " (the scratch holds " +
                                           std::to_string(kScratchBytes >> 20) + " MiB)");
  if (!row_ptr || !n_edges || !max_indegree)
    return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count_noalloc: null argument");
  *n_edges = 0;
  *max_indegree = 0;
  const long long rows = (long long)n_pre + 1;
  const int nchunk = (int)((rows + 1023) / 1024);
  const size_t off_csum = al256(sizeof(long long) * (size_t)rows);
  const size_t off_col = off_csum + al256(sizeof(long long) * (size_t)nchunk);
  const size_t off_max = off_col + al256(sizeof(int32_t) * (size_t)std::max(1, n_post));
  if (off_max + 256 > kScratchBytes)  // cannot happen within the limits above
    return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count_noalloc: scratch too small");
  int dev = 0;
  CCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return bs_internal_fail(BS_EINVAL, "bs_conn_bernoulli_count_noalloc: device index >= 64");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {  // once per device, with that device current
      void* ptr = nullptr;
      CCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = ptr;
    }
    base = (unsigned char*)g_scratch_ptr[dev];
  }
  long long* cnt = (long long*)base;
  long long* csum = (long long*)(base + off_csum);
  int32_t* col = (int32_t*)(base + off_col);
  int32_t* colmax = (int32_t*)(base + off_max);
  uint32_t k0, k1;
  conn_key(seed, k0, k1);
  const uint64_t thr = bern_threshold(p);
  // the scratch keeps the last call's contents: clear what must start at zero
  // (fill kernels: the row counts cnt[0 .. n_pre) are all written below)
  const unsigned fb = (unsigned)std::min<long long>(1024, (std::max(1, n_post) + 255) / 256);
  hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(256), 0, 0, col, (long long)std::max(1, n_post), 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, 0, colmax, 1ll, 0);
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, 0, cnt + n_pre, 1ll, 0ll);
  CCHK(hipGetLastError());
  if (n_pre > 0 && n_post > 0) {
    hipLaunchKernelGGL(k_conn_bern_count, dim3((unsigned)n_pre), dim3(256), 0, 0, k0, k1, (int)n_post, thr,
                       exclude_self ? 1 : 0, cnt, col);
  } else if (n_pre > 0) {
    const unsigned rb = (unsigned)std::min<long long>(1024, (n_pre + 255) / 256);
    hipLaunchKernelGGL(k_fill_i64, dim3(rb), dim3(256), 0, 0, cnt, (long long)n_pre, 0ll);
  }
  hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)nchunk), dim3(256), 0, 0, (const long long*)cnt, rows,
                     (long long*)row_ptr, csum);
  hipLaunchKernelGGL(k_scan_csum, dim3(1), dim3(256), 0, 0, csum, nchunk);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nchunk), dim3(256), 0, 0, (long long*)row_ptr, rows,
                     (const long long*)csum);
  if (n_post > 0)
    hipLaunchKernelGGL(k_max_i32, dim3(fb), dim3(256), 0, 0, (const int32_t*)col, (long long)n_post, colmax);
  CCHK(hipGetLastError());
  long long total = 0;
  int mx = 0;
  CCHK(hipMemcpy(&total, row_ptr + n_pre, sizeof(long long), hipMemcpyDeviceToHost));
  CCHK(hipMemcpy(&mx, colmax, sizeof(int), hipMemcpyDeviceToHost));
  *n_edges = total;
  *max_indegree = mx;
  return BS_OK;
}

}  // extern "C"
