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
#include <string>

#include "../../../include/brainspike_hip.h"
#include "bs_math.h"

int bs_internal_fail(int code, const std::string& msg);

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

}  // extern "C"
