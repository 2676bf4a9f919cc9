// bs_bptt.hip -- surrogate-gradient BPTT of a recurrent iaf_psc_exp network on
// gfx950: forward windows, checkpoint recomputation, reverse-time adjoint,
// transposed delayed delivery and deterministic weight gradients.
//
// Model and learner semantics (reference kernel path):
//   iaf_psc_exp.step       _kernel/models/iaf_psc_exp.py:120-151 (V held while
//                          refractory, PSCs decay then take this step's events)
//   SpikeRule.check / end  _kernel/spiking.py:364-488 (margin >= 0 and not
//                          refractory; reset to V_reset; counter set to
//                          t_ref/dt + 1 then decremented at the step end)
//   count_ste              _kernel/spiking.py:65-79 (hard count forward,
//                          surrogate derivative backward)
//   detach_reset = True    bptt.py:276 (no gradient through the reset)
//   rate_mse/membrane_mse  _kernel/learn/bptt.py:170-215
//   checkpointed_scan      _kernel/learn/bptt.py:437-458
//
// Layout in HBM (b = trial, j = neuron, N = n_exc + n_inh, NW = ceil(N/64)):
//   spike bitmaps  uint64 [T][B][NW]     written by the forward windows
//   checkpoints    V, I_ex, I_in, r [T/C][B][N]
//   act            float [C][B][N]       dV/dV surrogate factor per step of the
//                                        segment being reversed (-1: refractory)
//   ibar_ex/in     float [T][B][N]       gradient of the loss w.r.t. the input
//                                        arriving at (step, trial, neuron)
//   sbar           float [W][B][N]       gradient w.r.t. the spikes of a window:
//                                        W_rec . ibar(n + d), one GEMM per window
// Every sum runs in a fixed order (no float atomics), so results repeat
// bit-for-bit from run to run.

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../../include/brainspike_hip.h"

int bs_internal_fail(int code, const std::string& msg);

namespace {

#define BCHK(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return bs_internal_fail(BS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RBCHK(expr)                                                                      \
  do {                                                                                   \
    rocblas_status s_ = (expr);                                                          \
    if (s_ != rocblas_status_success)                                                    \
      return bs_internal_fail(BS_EHIP, std::string(#expr) + ": rocblas status " +        \
                                           std::to_string((int)s_));                     \
  } while (0)

constexpr int kTile = 256;       // neurons per block
constexpr int kMaxWords = 128;   // N <= 8192
constexpr int kMaxIn = 2048;     // input sources

struct Consts {
  float E_L, V_th, V_reset, I_e, P22, P21ex, P21in, P11ex, P11in, P20, span, alpha;
  int reset_to;
};

struct FwdArgs {
  Consts c;
  int N, NW, n_in, n_exc, B, T, d, d_in, C;
  int n0, nsteps;  // steps [n0, n0 + nsteps)
  int mode;        // 0: forward pass; 1: recompute a segment from its checkpoint
  int seg0;
  int loss;
  const float* w_in;
  const float* w_rec;
  const uint8_t* cnt;
  unsigned long long* bm;
  float *sv, *se, *si;
  int* sr;
  float *ck_v, *ck_e, *ck_i;
  int* ck_r;
  float* counts;
  double* lpart;
  const float* target;
  float* act;
  float* vbuf;
};

// Fast sigmoid (SuperSpike) derivative on the margin m = (V - V_th) / (V_th - V_reset):
// 1 / (alpha |m| + 1)^2, times dm/dV.
__device__ __forceinline__ float surrogate_dv(float m, const Consts& c) {
  const float q = c.alpha * fabsf(m) + 1.0f;
  return 1.0f / (q * q) / c.span;
}

__global__ __launch_bounds__(kTile) void k_bptt_fwd(FwdArgs a) {
  __shared__ unsigned long long bml[kMaxWords];
  __shared__ unsigned char cl[kMaxIn];
  const int tid = (int)threadIdx.x, b = (int)blockIdx.y;
  const int j = (int)blockIdx.x * kTile + tid;
  const bool on = j < a.N;
  const size_t BN = (size_t)a.B * a.N;
  const size_t bj = (size_t)b * a.N + (on ? j : 0);
  const Consts& c = a.c;
  float V = 0.f, ie = 0.f, ii = 0.f, cnt = 0.f;
  int r = 0;
  double lp = 0.0;
  if (on) {
    if (a.mode == 1) {
      const size_t o = (size_t)(a.seg0 / a.C) * BN + bj;
      V = a.ck_v[o];
      ie = a.ck_e[o];
      ii = a.ck_i[o];
      r = a.ck_r[o];
    } else {
      V = a.sv[bj];
      ie = a.se[bj];
      ii = a.si[bj];
      r = a.sr[bj];
      cnt = a.counts[bj];
      lp = a.lpart[bj];
    }
  }
  const float tgt = on ? a.target[j] : 0.f;
  for (int k = 0; k < a.nsteps; ++k) {
    const int n = a.n0 + k;
    if (a.mode == 0 && on && n % a.C == 0) {
      const size_t o = (size_t)(n / a.C) * BN + bj;
      a.ck_v[o] = V;
      a.ck_e[o] = ie;
      a.ck_i[o] = ii;
      a.ck_r[o] = r;
    }
    const int s_in = n - a.d_in, s_rec = n - a.d;
    __syncthreads();
    for (int t = tid; t < a.n_in; t += kTile)
      cl[t] = s_in >= 0 ? a.cnt[((size_t)s_in * a.B + b) * a.n_in + t] : (unsigned char)0;
    for (int t = tid; t < a.NW; t += kTile)
      bml[t] = s_rec >= 0 ? a.bm[((size_t)s_rec * a.B + b) * a.NW + t] : 0ull;
    __syncthreads();
    // inputs arriving at n: sources ascending, then presynaptic neurons ascending
    float xe = 0.f, xi = 0.f;
    if (on) {
      for (int i = 0; i < a.n_in; ++i) {
        const unsigned q = cl[i];
        if (q) xe = xe + (float)q * a.w_in[(size_t)i * a.N + j];
      }
      for (int w = 0; w < a.NW; ++w) {
        unsigned long long m = bml[w];
        while (m) {
          const int p = w * 64 + __builtin_ctzll(m);
          m &= m - 1;
          const float wv = a.w_rec[(size_t)p * a.N + j];
          if (p < a.n_exc) xe = xe + wv; else xi = xi + wv;
        }
      }
    }
    // the neuron update (iaf_psc_exp.py:136-151), then the spike rule
    const bool ref = r > 0;
    float y = V - c.E_L;
    const float y_new = ((y * c.P22 + ie * c.P21ex) + ii * c.P21in) + c.I_e * c.P20;
    if (!ref) y = y_new;
    ie = ie * c.P11ex + xe;
    ii = ii * c.P11in + xi;
    V = y + c.E_L;
    const float m = (V - c.V_th) / (c.V_th - c.V_reset);
    const bool hard = on && !ref && m >= 0.f;
    if (a.mode == 1 && on) {
      const size_t o = (size_t)(n - a.seg0) * BN + bj;
      a.act[o] = ref ? -1.0f : surrogate_dv(m, c);
    }
    if (hard) {
      V = c.V_reset;
      r = c.reset_to;
    }
    r = r > 0 ? r - 1 : 0;
    if (a.mode == 1 && on && a.vbuf) a.vbuf[(size_t)(n - a.seg0) * BN + bj] = V;
    if (a.mode == 0) {
      const unsigned long long mask = __ballot(hard);
      if ((tid & 63) == 0 && j < a.N) a.bm[((size_t)n * a.B + b) * a.NW + j / 64] = mask;
      if (hard) cnt += 1.0f;
      if (a.loss == BS_LOSS_MEMBRANE_MSE) {
        const double e = (double)V - (double)tgt;
        lp += e * e;
      }
    }
  }
  if (a.mode == 0 && on) {
    a.sv[bj] = V;
    a.se[bj] = ie;
    a.si[bj] = ii;
    a.sr[bj] = r;
    a.counts[bj] = cnt;
    a.lpart[bj] = lp;
  }
}

__global__ __launch_bounds__(kTile) void k_bptt_init(int B, int N, float E_L, const float* v0, float* sv,
                                                     float* se, float* si, int* sr, float* counts,
                                                     double* lpart, float* ay, float* ae, float* ai) {
  const size_t i = (size_t)blockIdx.x * kTile + threadIdx.x;
  if (i >= (size_t)B * N) return;
  if (sv) {
    sv[i] = v0 ? v0[i] : E_L;
    se[i] = 0.f;
    si[i] = 0.f;
    sr[i] = 0;
    counts[i] = 0.f;
    lpart[i] = 0.0;
  }
  if (ay) {
    ay[i] = 0.f;
    ae[i] = 0.f;
    ai[i] = 0.f;
  }
}

// Loss of the episode and, for rate_mse, d loss / d spike (constant over steps):
// one block, fixed summation order.
__global__ __launch_bounds__(kTile) void k_bptt_loss(int B, int N, int loss, double inv_T_s,
                                                     double inv_count, const float* counts,
                                                     const double* lpart, const float* target,
                                                     float* grate, double* out) {
  __shared__ double red[kTile];
  const int tid = (int)threadIdx.x;
  double acc = 0.0;
  for (size_t i = tid; i < (size_t)B * N; i += kTile) {
    if (loss == BS_LOSS_RATE_MSE) {
      const double rate = (double)counts[i] * inv_T_s;
      const double e = rate - (double)target[i % N];
      acc += e * e;
      grate[i] = (float)(2.0 * e * inv_count * inv_T_s);
    } else {
      acc += lpart[i];
    }
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = kTile / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) *out = red[0] * inv_count;
}

struct BwdArgs {
  Consts c;
  int N, NW, B, T;
  int n0, nsteps, seg0, loss;
  const unsigned long long* bm;
  const float* act;
  const float* vbuf;
  const float* sbar;
  const float* grate;
  const float* target;
  float mscale;
  float *ay, *ae, *ai;
  float *ibex, *ibin;
};

// Reverse-time adjoint of one window for one (trial, neuron tile). The carry
// holds d loss / d (V, I_ex, I_in) after step n; the input arriving at n
// enters I_ex/I_in at n, so its gradient is the carried I gradient.
__global__ __launch_bounds__(kTile) void k_bptt_bwd(BwdArgs a) {
  const int tid = (int)threadIdx.x, b = (int)blockIdx.y;
  const int j = (int)blockIdx.x * kTile + tid;
  if (j >= a.N) return;
  const size_t BN = (size_t)a.B * a.N;
  const size_t bj = (size_t)b * a.N + j;
  const Consts& c = a.c;
  float gv = a.ay[bj], ge = a.ae[bj], gi = a.ai[bj];
  const float g_rate = a.loss == BS_LOSS_RATE_MSE ? a.grate[bj] : 0.f;
  const float tgt = a.target[j];
  for (int k = a.nsteps - 1; k >= 0; --k) {
    const int n = a.n0 + k;
    const size_t o = (size_t)(n - a.seg0) * BN + bj;
    const float act = a.act[o];
    const bool ref = act < 0.f;
    const float ds = ref ? 0.f : act;
    const bool hard = (a.bm[((size_t)n * a.B + b) * a.NW + j / 64] >> (j & 63)) & 1ull;
    if (a.loss == BS_LOSS_MEMBRANE_MSE) gv = gv + a.mscale * 2.0f * (a.vbuf[o] - tgt);
    const float sb = a.sbar[((size_t)k * a.B + b) * a.N + j] + g_rate;
    a.ibex[(size_t)n * BN + bj] = ge;
    a.ibin[(size_t)n * BN + bj] = gi;
    const float gp = (hard ? 0.f : gv) + sb * ds;  // d loss / d V before the reset
    if (ref) {
      gv = gp;
      ge = ge * c.P11ex;
      gi = gi * c.P11in;
    } else {
      gv = gp * c.P22;
      ge = ge * c.P11ex + gp * c.P21ex;
      gi = gi * c.P11in + gp * c.P21in;
    }
  }
  a.ay[bj] = gv;
  a.ae[bj] = ge;
  a.ai[bj] = gi;
}

// d loss / d w_rec[p][:] = sum over the spikes of p at step s of
// ibar_{ex|in}(s + d)[:], in ascending (step, trial) order; one block per p.
__global__ __launch_bounds__(kTile) void k_bptt_wgrad_rec(int N, int NW, int B, int T, int d, int n_exc,
                                                          const unsigned long long* bm,
                                                          const float* ibex, const float* ibin,
                                                          float* g) {
  __shared__ int ev[kTile];
  __shared__ int wcount[kTile / 64];
  __shared__ int nev;
  const int p = (int)blockIdx.x, tid = (int)threadIdx.x;
  const float* ib = p < n_exc ? ibex : ibin;
  const int wi = p / 64, bit = p & 63;
  const long long npairs = (long long)(T - d) * B;  // arrival steps d..T-1
  for (int c0 = 0; c0 < N; c0 += 16 * kTile) {
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (long long base = 0; base < npairs; base += kTile) {
      const long long e = base + tid;
      bool hit = false;
      if (e < npairs) {
        const long long s = e / B;  // spike step (arrival s + d)
        const int bb = (int)(e % B);
        hit = (bm[((size_t)s * B + bb) * NW + wi] >> bit) & 1ull;
      }
      __syncthreads();
      // compact in ascending pair order: wave prefix via ballot, waves in order
      const unsigned long long bal = __ballot(hit);
      const int lane = tid & 63, wv = tid >> 6;
      if (lane == 0) wcount[wv] = __popcll(bal);
      __syncthreads();
      int off = 0;
      for (int w = 0; w < wv; ++w) off += wcount[w];
      if (hit) ev[off + __popcll(bal & ((1ull << lane) - 1ull))] = (int)(e - base);
      if (tid == 0) {
        int t = 0;
        for (int w = 0; w < kTile / 64; ++w) t += wcount[w];
        nev = t;
      }
      __syncthreads();
      const int m = nev;
      for (int k = 0; k < m; ++k) {
        const long long e2 = base + ev[k];
        const long long s = e2 / B;
        const int bb = (int)(e2 % B);
        const float* row = ib + ((size_t)(s + d) * B + bb) * N;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int col = c0 + q * kTile + tid;
          if (col < N) acc[q] = acc[q] + row[col];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int col = c0 + q * kTile + tid;
      if (col < N) g[(size_t)p * N + col] = acc[q];
    }
  }
}

// d loss / d w_in[i][:] = sum over (step, trial) of count_i(s) * ibar_ex(s + d_in)[:].
__global__ __launch_bounds__(kTile) void k_bptt_wgrad_in(int N, int B, int T, int d_in, int n_in,
                                                         const uint8_t* cnt, const float* ibex,
                                                         float* g) {
  __shared__ int ev[kTile];
  __shared__ unsigned char evc[kTile];
  __shared__ int wcount[kTile / 64];
  __shared__ int nev;
  const int i = (int)blockIdx.x, tid = (int)threadIdx.x;
  const long long npairs = (long long)(T - d_in) * B;
  for (int c0 = 0; c0 < N; c0 += 16 * kTile) {
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (long long base = 0; base < npairs; base += kTile) {
      const long long e = base + tid;
      unsigned char cv = 0;
      if (e < npairs) {
        const long long s = e / B;
        const int bb = (int)(e % B);
        cv = cnt[((size_t)s * B + bb) * n_in + i];
      }
      const bool hit = cv != 0;
      __syncthreads();
      const unsigned long long bal = __ballot(hit);
      const int lane = tid & 63, wv = tid >> 6;
      if (lane == 0) wcount[wv] = __popcll(bal);
      __syncthreads();
      int off = 0;
      for (int w = 0; w < wv; ++w) off += wcount[w];
      if (hit) {
        const int slot = off + __popcll(bal & ((1ull << lane) - 1ull));
        ev[slot] = (int)(e - base);
        evc[slot] = cv;
      }
      if (tid == 0) {
        int t = 0;
        for (int w = 0; w < kTile / 64; ++w) t += wcount[w];
        nev = t;
      }
      __syncthreads();
      const int m = nev;
      for (int k = 0; k < m; ++k) {
        const long long e2 = base + ev[k];
        const long long s = e2 / B;
        const int bb = (int)(e2 % B);
        const float cf = (float)evc[k];
        const float* row = ibex + ((size_t)(s + d_in) * B + bb) * N;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int col = c0 + q * kTile + tid;
          if (col < N) acc[q] = acc[q] + cf * row[col];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int col = c0 + q * kTile + tid;
      if (col < N) g[(size_t)i * N + col] = acc[q];
    }
  }
}

}  // namespace

struct bs_bptt {
  bs_bptt_desc d;
  Consts c;
  int N = 0, NW = 0, W = 0, nck = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::vector<void*> owned;
  unsigned long long* bm = nullptr;
  float *sv = nullptr, *se = nullptr, *si = nullptr;
  int* sr = nullptr;
  float *ck_v = nullptr, *ck_e = nullptr, *ck_i = nullptr;
  int* ck_r = nullptr;
  float* counts = nullptr;
  double* lpart = nullptr;
  float* grate = nullptr;
  double* d_loss = nullptr;
  float* act = nullptr;
  float* vbuf = nullptr;
  float *ibex = nullptr, *ibin = nullptr;
  float* sbar = nullptr;
  float *ay = nullptr, *ae = nullptr, *ai = nullptr;
  // arrays of the last forward (owned by the caller)
  const float* w_in = nullptr;
  const float* w_rec = nullptr;
  const uint8_t* cnt = nullptr;
  const float* target = nullptr;
  bool have_fwd = false;
};

namespace {

template <typename X>
int bmalloc(bs_bptt* h, X** p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&q, n * sizeof(X)) != hipSuccess) {
    (void)hipGetLastError();
    return bs_internal_fail(BS_ENOMEM, "bs_bptt_create: hipMalloc of " +
                                           std::to_string(n * sizeof(X)) + " bytes failed");
  }
  h->owned.push_back(q);
  *p = (X*)q;
  return BS_OK;
}

void free_bptt(bs_bptt* h) {
  if (!h) return;
  for (void* p : h->owned) (void)hipFree(p);
  if (h->blas) (void)rocblas_destroy_handle(h->blas);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

FwdArgs fwd_args(bs_bptt* h) {
  FwdArgs a{};
  a.c = h->c;
  a.N = h->N;
  a.NW = h->NW;
  a.n_in = h->d.n_in;
  a.n_exc = h->d.n_exc;
  a.B = h->d.batch;
  a.T = h->d.n_steps;
  a.d = h->d.delay_steps;
  a.d_in = h->d.delay_in_steps;
  a.C = h->d.checkpoint;
  a.loss = h->d.loss;
  a.w_in = h->w_in;
  a.w_rec = h->w_rec;
  a.cnt = h->cnt;
  a.bm = h->bm;
  a.sv = h->sv;
  a.se = h->se;
  a.si = h->si;
  a.sr = h->sr;
  a.ck_v = h->ck_v;
  a.ck_e = h->ck_e;
  a.ck_i = h->ck_i;
  a.ck_r = h->ck_r;
  a.counts = h->counts;
  a.lpart = h->lpart;
  a.target = h->target;
  a.act = h->act;
  a.vbuf = h->vbuf;
  return a;
}

}  // namespace

int bs_bptt_create(const bs_bptt_desc* d, bs_bptt** out) {
  if (!d || !out) return bs_internal_fail(BS_EINVAL, "bs_bptt_create: null argument");
  if (d->abi_version != BS_ABI_VERSION)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: descriptor ABI version mismatch");
  const long long N = (long long)d->n_exc + d->n_inh;
  if (d->n_exc < 0 || d->n_inh < 0 || N < 1 || N > 64 * kMaxWords)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: need 1 <= n_exc + n_inh <= 8192");
  if (d->n_in < 0 || d->n_in > kMaxIn)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: need 0 <= n_in <= 2048");
  if (d->batch < 1 || d->batch > 65535 || d->n_steps < 1)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: batch and n_steps must be >= 1");
  if (d->delay_steps < 1 || d->delay_in_steps < 1)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: delays must be >= 1 step");
  if (d->checkpoint < 1) return bs_internal_fail(BS_EINVAL, "bs_bptt_create: checkpoint must be >= 1");
  if (d->loss != BS_LOSS_RATE_MSE && d->loss != BS_LOSS_MEMBRANE_MSE)
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: unknown loss");
  if (!(d->V_th > d->V_reset))
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: V_th must exceed V_reset");
  if (d->ref_steps < 0) return bs_internal_fail(BS_EINVAL, "bs_bptt_create: ref_steps must be >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return bs_internal_fail(BS_ENODEV, "bs_bptt_create: no HIP device");
  }
  bs_bptt* h = new bs_bptt();
  h->d = *d;
  h->N = (int)N;
  h->NW = (int)((N + 63) / 64);
  h->W = std::min(d->delay_steps, 64);
  const int T = d->n_steps, B = d->batch, C = d->checkpoint;
  h->nck = (T + C - 1) / C;
  Consts& c = h->c;
  c.E_L = (float)d->E_L;
  c.V_th = (float)d->V_th;
  c.V_reset = (float)d->V_reset;
  c.I_e = (float)d->I_e;
  c.P22 = (float)d->P22;
  c.P21ex = (float)d->P21ex;
  c.P21in = (float)d->P21in;
  c.P11ex = (float)d->P11ex;
  c.P11in = (float)d->P11in;
  c.P20 = (float)d->P20;
  c.span = (float)d->V_th - (float)d->V_reset;
  c.alpha = (float)d->surrogate_alpha;
  c.reset_to = d->ref_steps > 0 ? d->ref_steps + 1 : 0;
  const size_t BN = (size_t)B * N;
  int rc = BS_OK;
#define BALLOC(p, n)                   \
  do {                                 \
    rc = bmalloc(h, &(p), (n));        \
    if (rc) {                          \
      free_bptt(h);                    \
      return rc;                       \
    }                                  \
  } while (0)
  BALLOC(h->bm, (size_t)T * B * h->NW);
  BALLOC(h->sv, BN);
  BALLOC(h->se, BN);
  BALLOC(h->si, BN);
  BALLOC(h->sr, BN);
  BALLOC(h->ck_v, (size_t)h->nck * BN);
  BALLOC(h->ck_e, (size_t)h->nck * BN);
  BALLOC(h->ck_i, (size_t)h->nck * BN);
  BALLOC(h->ck_r, (size_t)h->nck * BN);
  BALLOC(h->counts, BN);
  BALLOC(h->lpart, BN);
  BALLOC(h->grate, BN);
  BALLOC(h->d_loss, 1);
  BALLOC(h->act, (size_t)std::min(C, T) * BN);
  if (d->loss == BS_LOSS_MEMBRANE_MSE) BALLOC(h->vbuf, (size_t)std::min(C, T) * BN);
  BALLOC(h->ibex, (size_t)T * BN);
  BALLOC(h->ibin, (size_t)T * BN);
  BALLOC(h->sbar, (size_t)h->W * BN);
  BALLOC(h->ay, BN);
  BALLOC(h->ae, BN);
  BALLOC(h->ai, BN);
#undef BALLOC
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    free_bptt(h);
    return bs_internal_fail(BS_EHIP, "bs_bptt_create: hipStreamCreate failed");
  }
  if (rocblas_create_handle(&h->blas) != rocblas_status_success ||
      rocblas_set_stream(h->blas, h->stream) != rocblas_status_success) {
    free_bptt(h);
    return bs_internal_fail(BS_EHIP, "bs_bptt_create: rocblas handle");
  }
  *out = h;
  return BS_OK;
}

int bs_bptt_destroy(bs_bptt* h) {
  free_bptt(h);
  return BS_OK;
}

int bs_bptt_forward(bs_bptt* h, const float* w_in, const float* w_rec, const uint8_t* in_counts,
                    const float* v0, const float* target, double* loss) {
  if (!h || !w_rec || !target || !loss || (h->d.n_in > 0 && (!w_in || !in_counts)))
    return bs_internal_fail(BS_EINVAL, "bs_bptt_forward: null argument");
  h->w_in = w_in;
  h->w_rec = w_rec;
  h->cnt = in_counts;
  h->target = target;
  const int B = h->d.batch, T = h->d.n_steps, N = h->N;
  const size_t BN = (size_t)B * N;
  const unsigned nb = (unsigned)((BN + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_bptt_init, dim3(nb), dim3(kTile), 0, h->stream, B, N, h->c.E_L, v0, h->sv,
                     h->se, h->si, h->sr, h->counts, h->lpart, (float*)nullptr, (float*)nullptr,
                     (float*)nullptr);
  FwdArgs a = fwd_args(h);
  a.mode = 0;
  const dim3 grid((unsigned)((N + kTile - 1) / kTile), (unsigned)B);
  for (int n0 = 0; n0 < T; n0 += h->W) {
    a.n0 = n0;
    a.nsteps = std::min(h->W, T - n0);
    hipLaunchKernelGGL(k_bptt_fwd, grid, dim3(kTile), 0, h->stream, a);
  }
  const double T_s = T * h->d.dt_ms * 1e-3;
  hipLaunchKernelGGL(k_bptt_loss, dim3(1), dim3(kTile), 0, h->stream, B, N, h->d.loss, 1.0 / T_s,
                     h->d.loss == BS_LOSS_RATE_MSE ? 1.0 / (double)BN : 1.0 / ((double)BN * T),
                     h->counts, h->lpart, target, h->grate, h->d_loss);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(loss, h->d_loss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  BCHK(hipStreamSynchronize(h->stream));
  h->have_fwd = true;
  return BS_OK;
}

int bs_bptt_backward(bs_bptt* h, float* g_in, float* g_rec) {
  if (!h || !g_rec || (h->d.n_in > 0 && !g_in))
    return bs_internal_fail(BS_EINVAL, "bs_bptt_backward: null argument");
  if (!h->have_fwd) return bs_internal_fail(BS_EINVAL, "bs_bptt_backward: no forward pass to reverse");
  const int B = h->d.batch, T = h->d.n_steps, N = h->N, C = h->d.checkpoint;
  const int d = h->d.delay_steps, nE = h->d.n_exc, nI = h->d.n_inh;
  const size_t BN = (size_t)B * N;
  const unsigned nb = (unsigned)((BN + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_bptt_init, dim3(nb), dim3(kTile), 0, h->stream, B, N, 0.f, (const float*)nullptr,
                     (float*)nullptr, (float*)nullptr, (float*)nullptr, (int*)nullptr, (float*)nullptr,
                     (double*)nullptr, h->ay, h->ae, h->ai);
  const dim3 grid((unsigned)((N + kTile - 1) / kTile), (unsigned)B);
  FwdArgs fa = fwd_args(h);
  fa.mode = 1;
  BwdArgs ba{};
  ba.c = h->c;
  ba.N = N;
  ba.NW = h->NW;
  ba.B = B;
  ba.T = T;
  ba.loss = h->d.loss;
  ba.bm = h->bm;
  ba.act = h->act;
  ba.vbuf = h->vbuf;
  ba.sbar = h->sbar;
  ba.grate = h->grate;
  ba.target = h->target;
  ba.mscale = (float)(1.0 / ((double)BN * T));
  ba.ay = h->ay;
  ba.ae = h->ae;
  ba.ai = h->ai;
  ba.ibex = h->ibex;
  ba.ibin = h->ibin;
  const float one = 1.0f, zero = 0.0f;
  for (int seg = h->nck - 1; seg >= 0; --seg) {
    const int s0 = seg * C, s1 = std::min(T, s0 + C);
    fa.seg0 = s0;
    fa.n0 = s0;
    fa.nsteps = s1 - s0;
    hipLaunchKernelGGL(k_bptt_fwd, grid, dim3(kTile), 0, h->stream, fa);
    // windows of W steps from the end of the segment
    for (int w1 = s1; w1 > s0; w1 -= h->W) {
      const int w0 = std::max(s0, w1 - h->W);
      const int nsteps = w1 - w0;
      // sbar rows k: spikes at w0 + k reach their targets at w0 + k + d
      const int kmax = std::max(0, std::min(nsteps, T - d - w0));
      if (kmax < nsteps)
        BCHK(hipMemsetAsync(h->sbar + (size_t)kmax * BN, 0, sizeof(float) * (nsteps - kmax) * BN,
                            h->stream));
      if (kmax > 0) {
        const int M = kmax * B;
        const size_t src = (size_t)(w0 + d) * BN;
        if (nE > 0)
          RBCHK(rocblas_sgemm(h->blas, rocblas_operation_transpose, rocblas_operation_none, nE, M, N,
                              &one, h->w_rec, N, h->ibex + src, N, &zero, h->sbar, N));
        if (nI > 0)
          RBCHK(rocblas_sgemm(h->blas, rocblas_operation_transpose, rocblas_operation_none, nI, M, N,
                              &one, h->w_rec + (size_t)nE * N, N, h->ibin + src, N, &zero,
                              h->sbar + nE, N));
      }
      ba.n0 = w0;
      ba.nsteps = nsteps;
      ba.seg0 = s0;
      hipLaunchKernelGGL(k_bptt_bwd, grid, dim3(kTile), 0, h->stream, ba);
    }
  }
  if (T > d)
    hipLaunchKernelGGL(k_bptt_wgrad_rec, dim3((unsigned)N), dim3(kTile), 0, h->stream, N, h->NW, B, T, d,
                       nE, (const unsigned long long*)h->bm, (const float*)h->ibex,
                       (const float*)h->ibin, g_rec);
  else
    BCHK(hipMemsetAsync(g_rec, 0, sizeof(float) * (size_t)N * N, h->stream));
  if (h->d.n_in > 0) {
    if (T > h->d.delay_in_steps)
      hipLaunchKernelGGL(k_bptt_wgrad_in, dim3((unsigned)h->d.n_in), dim3(kTile), 0, h->stream, N, B, T,
                         h->d.delay_in_steps, h->d.n_in, h->cnt, (const float*)h->ibex, g_in);
    else
      BCHK(hipMemsetAsync(g_in, 0, sizeof(float) * (size_t)h->d.n_in * N, h->stream));
  }
  BCHK(hipGetLastError());
  BCHK(hipStreamSynchronize(h->stream));
  return BS_OK;
}

int bs_bptt_spikes(bs_bptt* h, uint64_t* bitmaps, float* counts) {
  if (!h) return bs_internal_fail(BS_EINVAL, "bs_bptt_spikes: null handle");
  if (!h->have_fwd) return bs_internal_fail(BS_EINVAL, "bs_bptt_spikes: no forward pass yet");
  const size_t BN = (size_t)h->d.batch * h->N;
  if (bitmaps)
    BCHK(hipMemcpy(bitmaps, h->bm, sizeof(uint64_t) * (size_t)h->d.n_steps * h->d.batch * h->NW,
                   hipMemcpyDeviceToHost));
  if (counts) BCHK(hipMemcpy(counts, h->counts, sizeof(float) * BN, hipMemcpyDeviceToHost));
  return BS_OK;
}
