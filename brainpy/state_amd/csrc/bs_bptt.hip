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
//   sbar           float [S][W][B][N]    gradient w.r.t. the spikes of a window:
//                                        W_rec . ibar(n + d), one GEMM per window
//                                        in S = kTdSplit partial sums over K
//   ibar operands  [T][B][Kp] per E/I    the GEMM's copies of ibar, K padded to a
//                                        multiple of 64 with zeros: bf16 hi and lo
//                                        parts (default) or float32
//   weight operand [N][Kp]               W_rec split (or copied) the same way,
//                                        once per forward
// The GEMM (k_bptt_tdeliver) is hand-written for the gfx950 matrix cores.
// Every sum runs in a fixed order (no float atomics), so results repeat
// bit-for-bit from run to run.

#include <hip/hip_runtime.h>
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
constexpr int kTile = 256;       // neurons per block
constexpr int kMaxWords = 128;   // N <= 8192
constexpr int kMaxIn = 512;      // input sources

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
  unsigned long long* bmT;  // [NW][T][B]
  float *sv, *se, *si;
  int* sr;
  float *ck_v, *ck_e, *ck_i;
  int* ck_r;
  float* counts;
  double* lpart;
  const float* target;
  float* act;
  float* vbuf;
  float *xin_e, *xin_i;  // summed inputs per (step, trial, neuron), or null
  int* err;
};

// Fast sigmoid (SuperSpike) derivative on the margin m = (V - V_th) / (V_th - V_reset):
// 1 / (alpha |m| + 1)^2, times dm/dV.
__device__ __forceinline__ float surrogate_dv(float m, const Consts& c) {
  const float q = c.alpha * fabsf(m) + 1.0f;
  return 1.0f / (q * q) / c.span;
}

// One step of iaf_psc_exp (iaf_psc_exp.py:136-151) on the summed inputs, then
// the spike rule; returns the margin m and whether the neuron was refractory.
// The forward pass and the segment replay both use it, so a replayed step is
// bit-identical to the forward one.
__device__ __forceinline__ bool iaf_exp_step(const Consts& c, float& V, float& ie, float& ii, int& r, float xe,
                                             float xi, bool on, float& mg) {
  const bool ref = r > 0;
  float y = V - c.E_L;
  const float y_new = ((y * c.P22 + ie * c.P21ex) + ii * c.P21in) + c.I_e * c.P20;
  if (!ref) y = y_new;
  ie = ie * c.P11ex + xe;
  ii = ii * c.P11in + xi;
  V = y + c.E_L;
  mg = (V - c.V_th) / (c.V_th - c.V_reset);
  const bool hard = on && !ref && mg >= 0.f;
  if (hard) {
    V = c.V_reset;
    r = c.reset_to;
  }
  r = r > 0 ? r - 1 : 0;
  return ref;
}

constexpr int kChunk = 8;      // steps whose events are compacted together
constexpr int kEvCap = 3072;   // events per chunk held in LDS (~28 KB of LDS per
                               // block keeps 1024 blocks resident on 256 CUs)
constexpr int kBatch = 16;     // event loads in flight per neuron

__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  return inc - v;
}

// Weight of event row r for neuron j: rows 0..n_in-1 are w_in, the rest w_rec.
__device__ __forceinline__ float ev_weight(const FwdArgs& a, int row, int j) {
  return row < a.n_in ? a.w_in[(size_t)row * a.N + j] : a.w_rec[(size_t)(row - a.n_in) * a.N + j];
}

// Forward steps of one (trial, neuron tile). The events arriving in a chunk
// of steps (inputs with a non-zero count, then spiking presynaptic neurons,
// both ascending) are compacted into LDS first; each step then sums them in
// that order while the next step's weights are already being loaded.
__global__ __launch_bounds__(kTile) void k_bptt_fwd(FwdArgs a) {
  __shared__ unsigned short ev_row[kEvCap];
  __shared__ unsigned char ev_sc[kEvCap];
  __shared__ unsigned char ev_k[kEvCap];
  __shared__ float xs[2][kChunk][kTile];  // summed inputs (ex, in) per step of the chunk
  __shared__ unsigned char cl[kChunk][kMaxIn];  // input counts of the chunk's steps
  __shared__ int ev_off[kChunk + 1];
  __shared__ int st_cnt[kChunk];
  const int tid = (int)threadIdx.x, b = (int)blockIdx.y;
  const int lane = tid & 63, wave = tid >> 6;
  const int j = (int)blockIdx.x * kTile + tid;
  const bool on = j < a.N;
  const int jj = on ? j : 0;
  const size_t BN = (size_t)a.B * a.N;
  const size_t bj = (size_t)b * a.N + jj;
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
  const int nex_rows = a.n_in + a.n_exc;  // rows below this feed I_syn_ex
  for (int c0 = 0; c0 < a.nsteps; c0 += kChunk) {
    const int nc = min(kChunk, a.nsteps - c0);
    // ---- count the events of every step of the chunk (one wave per step):
    // the step's spike words stay in registers, its input counts in LDS
    static_assert(kChunk <= 2 * (kTile / 64), "two steps per wave");
    unsigned long long wr[2][2];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int k = wave + j * (kTile / 64);
      wr[j][0] = wr[j][1] = 0ull;
      if (k < nc) {
        const int n = a.n0 + c0 + k;
        const int s_in = n - a.d_in, s_rec = n - a.d;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int w = lane + 64 * t;
          if (s_rec >= 0 && w < a.NW) wr[j][t] = a.bm[((size_t)s_rec * a.B + b) * a.NW + w];
        }
        int m = __popcll(wr[j][0]) + __popcll(wr[j][1]);
        for (int i = lane; i < a.n_in; i += 64) {
          const unsigned char q = s_in >= 0 ? a.cnt[((size_t)s_in * a.B + b) * a.n_in + i] : (unsigned char)0;
          cl[k][i] = q;
          m += q != 0;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
        if (lane == 0) st_cnt[k] = m;
      }
    }
    __syncthreads();
    if (tid == 0) {
      int t = 0;
      for (int k = 0; k < nc; ++k) {
        ev_off[k] = t;
        t += st_cnt[k];
      }
      ev_off[nc] = t;
    }
    __syncthreads();
    if (ev_off[nc] > kEvCap) {  // more events than LDS holds: report, leave state untouched
      if (tid == 0) atomicOr(a.err, 1);
      return;
    }
    // ---- write the lists: inputs ascending, then presynaptic neurons ascending
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int k = wave + j * (kTile / 64);
      if (k >= nc) continue;
      int pos = ev_off[k];
      for (int i0 = 0; i0 < a.n_in; i0 += 64) {
        const int i = i0 + lane;
        const unsigned q = i < a.n_in ? cl[k][i] : 0u;
        const unsigned long long bal = __ballot(q != 0);
        if (q) {
          const int slot = pos + __popcll(bal & ((1ull << lane) - 1ull));
          ev_row[slot] = (unsigned short)i;
          ev_sc[slot] = (unsigned char)q;
          ev_k[slot] = (unsigned char)k;
        }
        pos += __popcll(bal);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        if (64 * t >= a.NW) break;
        unsigned long long word = wr[j][t];
        const int pc = __popcll(word);
        int slot = pos + wave_excl_scan(pc, lane);
        while (word) {
          const int bit = __builtin_ctzll(word);
          word &= word - 1;
          ev_row[slot] = (unsigned short)(a.n_in + (lane + 64 * t) * 64 + bit);
          ev_sc[slot] = 1;
          ev_k[slot] = (unsigned char)k;
          ++slot;
        }
        pos = __shfl(slot, 63, 64);  // end of lane 63's bits
      }
    }
    __syncthreads();
    // ---- gather: every event of the chunk in list order (per step: inputs,
    // then presynaptic neurons, ascending), kBatch loads in flight. The list
    // is ordered by step, so each step's two sums run in registers (the same
    // additions in the same order as summing in place) and are written to
    // this neuron's LDS slots when the list moves to the next step.
    for (int k = 0; k < nc; ++k) {
      xs[0][k][tid] = 0.f;
      xs[1][k][tid] = 0.f;
    }
    const int E = ev_off[nc];
    float ae = 0.f, ai = 0.f;
    int kcur = 0;
    for (int e0 = 0; e0 < E; e0 += kBatch) {
      float v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) v[u] = e0 + u < E ? ev_weight(a, ev_row[e0 + u], jj) : 0.f;
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (e0 + u < E) {
          const int k = ev_k[e0 + u];
          if (k != kcur) {  // block-uniform: the list is shared
            xs[0][kcur][tid] = ae;
            xs[1][kcur][tid] = ai;
            ae = ai = 0.f;
            kcur = k;
          }
          const int row = ev_row[e0 + u];
          const float x = row < a.n_in ? (float)ev_sc[e0 + u] * v[u] : v[u];
          if (row < nex_rows) ae = ae + x;
          else ai = ai + x;
        }
      }
    }
    xs[0][kcur][tid] = ae;
    xs[1][kcur][tid] = ai;
    // ---- the steps
    for (int k = 0; k < nc; ++k) {
      const int n = a.n0 + c0 + k;
      if (a.mode == 0 && on && n % a.C == 0) {
        const size_t o = (size_t)(n / a.C) * BN + bj;
        a.ck_v[o] = V;
        a.ck_e[o] = ie;
        a.ck_i[o] = ii;
        a.ck_r[o] = r;
      }
      const float xe = xs[0][k][tid], xi = xs[1][k][tid];
      if (a.mode == 0 && on && a.xin_e) {
        a.xin_e[(size_t)n * BN + bj] = xe;
        a.xin_i[(size_t)n * BN + bj] = xi;
      }
      float mg;
      const bool ref = iaf_exp_step(c, V, ie, ii, r, xe, xi, on, mg);
      const bool hard = on && !ref && mg >= 0.f;
      if (a.mode == 1 && on) {
        const size_t o = (size_t)(n - a.seg0) * BN + bj;
        a.act[o] = ref ? -1.0f : surrogate_dv(mg, c);
      }
      if (a.mode == 1 && on && a.vbuf) a.vbuf[(size_t)(n - a.seg0) * BN + bj] = V;
      if (a.mode == 0) {
        const unsigned long long mask = __ballot(hard);
        if (lane == 0 && j < a.N) {
          a.bm[((size_t)n * a.B + b) * a.NW + j / 64] = mask;
          a.bmT[((size_t)(j / 64) * a.T + n) * a.B + b] = mask;
        }
        if (hard) cnt += 1.0f;
        if (a.loss == BS_LOSS_MEMBRANE_MSE) {
          const double e = (double)V - (double)tgt;
          lp += e * e;
        }
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

// Replay of a checkpoint segment from the stored summed inputs: the same
// neuron steps as the forward pass (bit-identical V), writing the surrogate
// factors (and V for the membrane loss) the reverse pass reads.
__global__ __launch_bounds__(kTile) void k_bptt_replay(FwdArgs a) {
  const int b = (int)blockIdx.y;
  const int j = (int)blockIdx.x * kTile + (int)threadIdx.x;
  if (j >= a.N) return;
  const size_t BN = (size_t)a.B * a.N;
  const size_t bj = (size_t)b * a.N + j;
  const Consts& c = a.c;
  const size_t o0 = (size_t)(a.seg0 / a.C) * BN + bj;
  float V = a.ck_v[o0], ie = a.ck_e[o0], ii = a.ck_i[o0];
  int r = a.ck_r[o0];
  for (int k = 0; k < a.nsteps; ++k) {
    const int n = a.n0 + k;
    const float xe = a.xin_e[(size_t)n * BN + bj], xi = a.xin_i[(size_t)n * BN + bj];
    float mg;
    const bool ref = iaf_exp_step(c, V, ie, ii, r, xe, xi, true, mg);
    const size_t o = (size_t)(n - a.seg0) * BN + bj;
    a.act[o] = ref ? -1.0f : surrogate_dv(mg, c);
    if (a.vbuf) a.vbuf[o] = V;
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

// Loss of the episode and, for rate_mse, d loss / d spike (constant over
// steps). Two fixed-order reduction levels: kLossBlocks blocks each sum one
// contiguous range, then one block sums the partials in block order.
constexpr int kLossBlocks = 256;

__global__ __launch_bounds__(kTile) void k_bptt_loss_part(int B, int N, int loss, double inv_T_s,
                                                          double inv_count, const float* counts,
                                                          const double* lpart, const float* target,
                                                          float* grate, double* part) {
  __shared__ double red[kTile];
  const int tid = (int)threadIdx.x;
  const size_t total = (size_t)B * N;
  const size_t per = (total + kLossBlocks - 1) / kLossBlocks;
  const size_t lo = (size_t)blockIdx.x * per, hi = std::min(total, lo + per);
  double acc = 0.0;
  for (size_t i = lo + tid; i < hi; i += kTile) {
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
  if (tid == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kTile) void k_bptt_loss_final(double inv_count, const double* part,
                                                           double* out) {
  __shared__ double red[kTile];
  const int tid = (int)threadIdx.x;
  red[tid] = tid < kLossBlocks ? part[tid] : 0.0;
  __syncthreads();
  for (int s = kTile / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) *out = red[0] * inv_count;
}

// bf16 round-to-nearest-even of a finite float, as raw bits.
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Operand rows of the transposed delivery: [rows][Kp], K padded with zeros.
// mode 0: the bf16 parts hi = bf16(x), lo = bf16(x - hi) in two arrays; mode 1:
// float32 copies.
__global__ __launch_bounds__(kTile) void k_bptt_prep_w(int N, int Kp, int mode, const float* w,
                                                       unsigned short* wh, unsigned short* wl, float* wf) {
  const int p = (int)blockIdx.y;
  for (int k = (int)blockIdx.x * kTile + (int)threadIdx.x; k < Kp; k += (int)gridDim.x * kTile) {
    const float x = k < N ? w[(size_t)p * N + k] : 0.f;
    const size_t o = (size_t)p * Kp + k;
    if (mode == 0) {
      const unsigned short hi = bf16_bits(x);
      wh[o] = hi;
      wl[o] = bf16_bits(x - bf16_float(hi));
    } else {
      wf[o] = x;
    }
  }
}

// ---- transposed delayed delivery on the matrix cores
// out[m][i] = sum_k IB_p[m][k] W[i][k] for the window's rows m = (step, trial)
// and presynaptic neurons i, p = 0 (excitatory rows, IB = ibar of I_ex) or 1
// (inhibitory rows, IB = ibar of I_in). Mode 0 computes each product as
// hi*hi + hi*lo + lo*hi of the bf16 parts (v_mfma_f32_32x32x16_bf16, float32
// accumulation: about float32 accuracy at the bf16 rate); mode 1 multiplies the
// float32 values (v_mfma_f32_32x32x2f32). Each output is summed in one fixed
// order, so results repeat bit for bit.
//
// Block: 4 waves (2 over m x 2 over i), tile 128 m x 128 i; a wave owns 64 m x
// 64 i (four 32x32 accumulator tiles), so each fragment read from LDS feeds two
// MFMA tiles. K advances in stages of 32: every stage of both operands is
// loaded (16-byte loads, 8 k per load) into registers one stage ahead and
// stored into the other half of a double-buffered LDS area (64 KiB: two blocks
// per CU). Blocks are dealt round-robin over the 8 XCDs (block b on XCD b % 8),
// so each XCD is given one contiguous range of (K range, i-tile, m-tile)
// triples: the blocks of an XCD share their weight rows and the window's ibar
// rows of one K range in its L2.
constexpr int kTdBM = 128, kTdBI = 128, kTdKC = 32;
// K is split into kTdSplit ranges, one block per (tile, range), each writing its
// own partial sum; the backward kernel adds the partials in range order. The
// benchmark window (640 rows x 4096 presynaptic rows) has 165 tiles: in thirds
// 495 blocks, about two per CU (in halves 330 blocks leave 74 CUs with two and
// the rest with one; measured per gradient step: halves 21.9 ms, thirds 20.7,
// quarters 22.7).
#ifndef BS_TD_SPLIT
#define BS_TD_SPLIT 3
#endif
constexpr int kTdSplit = BS_TD_SPLIT;

struct TdArgs {
  const void* w0;         // mode 0: W hi parts, mode 1: W float32   [N][Kp]
  const void* w1;         // mode 0: W lo parts
  const void* b0[2];      // per part: IB hi parts / float32, rows m [M][Kp]
  const void* b1[2];      // per part: IB lo parts
  float* out;             // [nks][M][N] partial sums, partial s at out + s * pstride
  size_t pstride;
  int M, N, Kp, n_exc, nks;
  int nti0, nti1, ntm;    // i-tiles of the excitatory and inhibitory rows, m-tiles
};

// LDS slot (16-byte unit) of group g of row r in a stage holding KB groups per
// row: the group index is XOR-ed with row bits so that staging stores (threads
// t..t+KB-1: one row's groups) and fragment reads (lanes: consecutive rows,
// one group) both spread over the 64 four-byte banks.
template <int KB>
__device__ __forceinline__ int td_slot(int r, int g) {
  constexpr int sh = KB == 4 ? 2 : 1;  // rows per 256 bytes: 4 (KB = 4), 2 (KB = 8)
  return r * KB + (g ^ ((r >> sh) & (KB - 1)));
}

__device__ __forceinline__ uint4 td_mask(uint4 v, unsigned m) {
  return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m);
}

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 td_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float td_f32x16;

template <int MODE>
__global__ __launch_bounds__(256) void k_bptt_tdeliver(TdArgs a) {
  constexpr int BM = kTdBM, BI = kTdBI, KC = kTdKC;
  constexpr int KB = MODE == 0 ? KC / 8 : KC / 4;   // 16-byte groups per row and array
  constexpr int NA = MODE == 0 ? 2 : 1;             // arrays per operand
  constexpr int NWL = BI * KB / 256, NBL = BM * KB / 256;  // 16-byte loads per thread and array
  constexpr int ESZ = MODE == 0 ? 2 : 4;            // element bytes
  constexpr int STG = NA * (BI + BM) * KB;          // 16-byte slots per stage
  __shared__ uint4 lds[2 * STG];
  const int tiles = (a.nti0 + a.nti1) * a.ntm;
  const int total = tiles * a.nks;
  const int per = (total + 7) / 8;
  const int L0 = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
  if (L0 >= total) return;
  // K range major, so an XCD's blocks share one range of both operands
  const int ks = L0 / tiles, L = L0 - ks * tiles;
  const int it = L / a.ntm, mt = L - it * a.ntm;
  const int nk_all = a.Kp / KC;
  const int kbeg = ks * nk_all / a.nks, kend = (ks + 1) * nk_all / a.nks;
  float* const out = a.out + (size_t)ks * a.pstride;
  const int part = it < a.nti0 ? 0 : 1;
  const int i0 = part == 0 ? it * BI : a.n_exc + (it - a.nti0) * BI;
  const int i1 = part == 0 ? a.n_exc : a.N;
  const int m0 = mt * BM;
  const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = wv & 1, wi = wv >> 1;
  constexpr int MT = BM / 64, IT = BI / 64;  // 32x32 accumulator tiles per wave along m and i
  const unsigned char* W0 = (const unsigned char*)a.w0;
  const unsigned char* W1 = (const unsigned char*)a.w1;
  const unsigned char* B0 = (const unsigned char*)a.b0[part];
  const unsigned char* B1 = (const unsigned char*)a.b1[part];
  const size_t rowb = (size_t)a.Kp * ESZ;
  // loads are unconditional (rows clamped into range) so that the compiler can
  // count them; rows past the tile's end are zeroed when stored
  uint4 rw[NA][NWL], rb[NA][NBL];
#define TD_LOAD(k0_)                                                                    \
  do {                                                                                  \
    _Pragma("unroll") for (int j = 0; j < NWL; ++j) {                                   \
      const int c = tid + 256 * j, r = c / KB, g = c % KB;                              \
      const size_t o = (size_t)min(i0 + r, i1 - 1) * rowb + ((size_t)(k0_) * ESZ + g * 16); \
      rw[0][j] = *(const uint4*)(W0 + o);                                               \
      if (NA > 1) rw[NA - 1][j] = *(const uint4*)(W1 + o);                              \
    }                                                                                   \
    _Pragma("unroll") for (int j = 0; j < NBL; ++j) {                                   \
      const int c = tid + 256 * j, r = c / KB, g = c % KB;                              \
      const size_t o = (size_t)min(m0 + r, a.M - 1) * rowb + ((size_t)(k0_) * ESZ + g * 16); \
      rb[0][j] = *(const uint4*)(B0 + o);                                               \
      if (NA > 1) rb[NA - 1][j] = *(const uint4*)(B1 + o);                              \
    }                                                                                   \
  } while (0)
#define TD_STORE(st_)                                                                   \
  do {                                                                                  \
    uint4* s_ = (st_);                                                                  \
    _Pragma("unroll") for (int j = 0; j < NWL; ++j) {                                   \
      const int c = tid + 256 * j, r = c / KB, g = c % KB;                              \
      const unsigned mk = i0 + r < i1 ? ~0u : 0u;                                       \
      _Pragma("unroll") for (int q = 0; q < NA; ++q)                                    \
        s_[q * BI * KB + td_slot<KB>(r, g)] = td_mask(rw[q][j], mk);                    \
    }                                                                                   \
    _Pragma("unroll") for (int j = 0; j < NBL; ++j) {                                   \
      const int c = tid + 256 * j, r = c / KB, g = c % KB;                              \
      const unsigned mk = m0 + r < a.M ? ~0u : 0u;                                      \
      _Pragma("unroll") for (int q = 0; q < NA; ++q)                                    \
        s_[NA * BI * KB + q * BM * KB + td_slot<KB>(r, g)] = td_mask(rb[q][j], mk);     \
    }                                                                                   \
  } while (0)
  td_f32x16 acc[MT][IT];
#pragma unroll
  for (int mt_ = 0; mt_ < MT; ++mt_)
#pragma unroll
    for (int t = 0; t < IT; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[mt_][t][r] = 0.f;
  const int nk = kend - kbeg;
  if (nk > 0) {
  TD_LOAD(kbeg * KC);
  TD_STORE(lds);
  }
  __syncthreads();
  const int rb_ = wm * (BM / 2) + (lane & 31);
  for (int k = 0; k < nk; ++k) {
    if (k + 1 < nk) TD_LOAD((kbeg + k + 1) * KC);
    const uint4* st = lds + (k & 1) * STG;
    const uint4* sW = st;
    const uint4* sB = st + NA * BI * KB;
    if constexpr (MODE == 0) {
#pragma unroll
      for (int s2 = 0; s2 < KC / 16; ++s2) {
        const int g = 2 * s2 + (lane >> 5);
        td_bf16x8 ah[MT], al[MT];
#pragma unroll
        for (int mt_ = 0; mt_ < MT; ++mt_) {
          ah[mt_] = __builtin_bit_cast(td_bf16x8, sB[td_slot<KB>(rb_ + 32 * mt_, g)]);
          al[mt_] = __builtin_bit_cast(td_bf16x8, sB[BM * KB + td_slot<KB>(rb_ + 32 * mt_, g)]);
        }
#pragma unroll
        for (int t = 0; t < IT; ++t) {
          const int r = wi * (BI / 2) + t * 32 + (lane & 31);
          const td_bf16x8 wh = __builtin_bit_cast(td_bf16x8, sW[td_slot<KB>(r, g)]);
          const td_bf16x8 wl = __builtin_bit_cast(td_bf16x8, sW[BI * KB + td_slot<KB>(r, g)]);
#pragma unroll
          for (int mt_ = 0; mt_ < MT; ++mt_) {
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt_], wh, acc[mt_][t], 0, 0, 0);
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt_], wl, acc[mt_][t], 0, 0, 0);
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mt_], wh, acc[mt_][t], 0, 0, 0);
          }
        }
      }
    } else {
      // lanes 0-31 hold k = 8 s2 + e, lanes 32-63 k = 8 s2 + 4 + e of their
      // rows' float4 groups; MFMA e (k depth 2) takes element e of both operands
#pragma unroll
      for (int s2 = 0; s2 < KC / 8; ++s2) {
        const int g = 2 * s2 + (lane >> 5);
        float4 av[MT];
#pragma unroll
        for (int mt_ = 0; mt_ < MT; ++mt_) av[mt_] = __builtin_bit_cast(float4, sB[td_slot<KB>(rb_ + 32 * mt_, g)]);
#pragma unroll
        for (int t = 0; t < IT; ++t) {
          const int r = wi * (BI / 2) + t * 32 + (lane & 31);
          const float4 wv4 = __builtin_bit_cast(float4, sW[td_slot<KB>(r, g)]);
#pragma unroll
          for (int mt_ = 0; mt_ < MT; ++mt_) {
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt_].x, wv4.x, acc[mt_][t], 0, 0, 0);
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt_].y, wv4.y, acc[mt_][t], 0, 0, 0);
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt_].z, wv4.z, acc[mt_][t], 0, 0, 0);
            acc[mt_][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt_].w, wv4.w, acc[mt_][t], 0, 0, 0);
          }
        }
      }
    }
    if (k + 1 < nk) TD_STORE(lds + ((k + 1) & 1) * STG);
    __syncthreads();
  }
#undef TD_LOAD
#undef TD_STORE
  // accumulator layout of a 32x32 tile: lane l holds column i = l % 32 and
  // rows m = 8 (r / 4) + 4 (l / 32) + r % 4 for r = 0..15
#pragma unroll
  for (int t = 0; t < IT; ++t) {
    const int i = i0 + wi * (BI / 2) + t * 32 + (lane & 31);
    if (i >= i1) continue;
#pragma unroll
    for (int mt_ = 0; mt_ < MT; ++mt_)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * (BM / 2) + 32 * mt_ + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m < a.M) out[(size_t)m * a.N + i] = acc[mt_][t][r];
      }
  }
}

struct BwdArgs {
  Consts c;
  int N, NW, B, T;
  int n0, nsteps, seg0, loss;
  const unsigned long long* bm;
  const float* act;
  const float* vbuf;
  const float* sbar;     // nsplit partial sums, partial s at sbar + s * pstride
  size_t pstride;
  int nsplit;
  const float* grate;
  const float* target;
  float mscale;
  float *ay, *ae, *ai;
  float *ibex, *ibin;
  int Kp, mode;                  // GEMM operand rows: Kp columns, mode 0 bf16 parts, 1 float32
  unsigned short *ibh[2], *ibl[2];  // mode 0: hi / lo parts of ibar (E, I) [T][B][Kp]
  float* ibf[2];                    // mode 1: float32 copies of ibar (E, I) [T][B][Kp]
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
    const size_t so = ((size_t)k * a.B + b) * a.N + j;
    float sb = a.sbar[so];
    for (int s = 1; s < a.nsplit; ++s) sb = sb + a.sbar[(size_t)s * a.pstride + so];
    sb = sb + g_rate;
    a.ibex[(size_t)n * BN + bj] = ge;
    a.ibin[(size_t)n * BN + bj] = gi;
    const size_t ro = ((size_t)n * a.B + b) * a.Kp + j;  // the GEMM's operand rows
    if (a.mode == 0) {
      const unsigned short he = bf16_bits(ge), hi_ = bf16_bits(gi);
      a.ibh[0][ro] = he;
      a.ibl[0][ro] = bf16_bits(ge - bf16_float(he));
      a.ibh[1][ro] = hi_;
      a.ibl[1][ro] = bf16_bits(gi - bf16_float(hi_));
    } else {
      a.ibf[0][ro] = ge;
      a.ibf[1][ro] = gi;
    }
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

// Weight gradients, one block per presynaptic row (a recurrent neuron p or an
// input source i): d loss / d w[row][:] = sum over the row's events at step s
// (trial b) of scale * ibar(s + delay)[b][:], in ascending (s, b) order. The
// events are found by scanning the row's column of the transposed spike
// bitmaps (or input counts), contiguous in (s, b).
struct WgArgs {
  int N, B, T, d, kind;  // kind 0: recurrent rows, 1: input rows
  int n_exc;
  const unsigned long long* bmT;  // [NW][T][B]
  const uint8_t* cntT;            // [n_in][T][B]
  const float* ibex;
  const float* ibin;
  float* g;
};

__global__ __launch_bounds__(kTile) void k_bptt_wgrad(WgArgs a) {
  __shared__ int ev[kTile];
  __shared__ float evs[kTile];
  __shared__ int wcount[kTile / 64];
  __shared__ int nev;
  const int row = (int)blockIdx.x, tid = (int)threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const size_t TB = (size_t)a.T * a.B;
  const float* ib = (a.kind == 0 && row >= a.n_exc) ? a.ibin : a.ibex;
  const unsigned long long* wcol = a.kind == 0 ? a.bmT + (size_t)(row / 64) * TB : nullptr;
  const uint8_t* ccol = a.kind == 1 ? a.cntT + (size_t)row * TB : nullptr;
  const int bit = row & 63;
  const long long npairs = (long long)(a.T - a.d) * a.B;
  for (int c0 = 0; c0 < a.N; c0 += 16 * kTile) {
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (long long base = 0; base < npairs; base += kTile) {
      const long long e = base + tid;
      float sc = 0.f;
      if (e < npairs) {
        if (a.kind == 0) sc = (float)((wcol[e] >> bit) & 1ull);
        else sc = (float)ccol[e];
      }
      const bool hit = sc != 0.f;
      __syncthreads();
      const unsigned long long bal = __ballot(hit);
      if (lane == 0) wcount[wv] = __popcll(bal);
      __syncthreads();
      int off = 0;
      for (int w = 0; w < wv; ++w) off += wcount[w];
      if (hit) {
        const int slot = off + __popcll(bal & ((1ull << lane) - 1ull));
        ev[slot] = (int)(e - base);
        evs[slot] = sc;
      }
      if (tid == 0) {
        int t = 0;
        for (int w = 0; w < kTile / 64; ++w) t += wcount[w];
        nev = t;
      }
      __syncthreads();
      const int m = nev;
      // two event rows in flight per pass (87 VGPRs, five waves per SIMD;
      // four rows took 119 VGPRs and four waves, 0.4 ms more per gradient step)
      for (int k = 0; k < m; k += 2) {
        float v[2][16];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const float* rp = nullptr;
          if (k + u < m) {
            const long long e2 = base + ev[k + u];
            const long long st = e2 / a.B;
            const int bb = (int)(e2 % a.B);
            rp = ib + ((size_t)(st + a.d) * a.B + bb) * a.N;
          }
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int col = c0 + q * kTile + tid;
            v[u][q] = (rp && col < a.N) ? rp[col] : 0.f;
          }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          if (k + u < m) {
            const float f = evs[k + u];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = acc[q] + f * v[u][q];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int col = c0 + q * kTile + tid;
      if (col < a.N) a.g[(size_t)row * a.N + col] = acc[q];
    }
  }
}

// counts [T][B][n_in] -> [n_in][T][B]
__global__ __launch_bounds__(kTile) void k_bptt_transpose_counts(int n_in, size_t TB, const uint8_t* cnt,
                                                                 uint8_t* cntT) {
  const size_t idx = (size_t)blockIdx.x * kTile + threadIdx.x;
  if (idx >= TB * n_in) return;
  const size_t sb = idx / n_in;
  const int i = (int)(idx % n_in);
  cntT[(size_t)i * TB + sb] = cnt[idx];
}

}  // namespace

struct bs_bptt {
  bs_bptt_desc d;
  Consts c;
  int N = 0, NW = 0, W = 0, nck = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> owned;
  unsigned long long* bm = nullptr;
  unsigned long long* bmT = nullptr;
  uint8_t* cntT = nullptr;
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
  float *xin_e = nullptr, *xin_i = nullptr;  // stored summed inputs (segment replay)
  int Kp = 0;                                           // GEMM operand row length
  unsigned short *wh = nullptr, *wl = nullptr;          // mode 0: W_rec parts [N][Kp]
  unsigned short *ibh[2] = {nullptr, nullptr}, *ibl[2] = {nullptr, nullptr};
  float* wf = nullptr;                                  // mode 1: W_rec [N][Kp]
  float* ibf[2] = {nullptr, nullptr};

  double* lpart_blocks = nullptr;
  float* sbar = nullptr;
  float *ay = nullptr, *ae = nullptr, *ai = nullptr;
  int* err = nullptr;
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
  a.bmT = h->bmT;
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
  a.xin_e = h->xin_e;
  a.xin_i = h->xin_i;
  a.err = h->err;
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
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: need 0 <= n_in <= 512");
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
  BALLOC(h->bmT, (size_t)T * B * h->NW);
  BALLOC(h->cntT, (size_t)T * B * std::max(1, d->n_in));
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
  // With replay, the forward pass also stores every step's summed inputs
  // (by default when they take at most 32 GiB): the reverse pass then replays
  // each checkpoint segment with the neuron steps alone instead of gathering
  // its events again.
  if (d->replay < -1 || d->replay > 1) {
    free_bptt(h);
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: replay must be -1, 0 or 1");
  }
  if (d->replay == 1 || (d->replay == -1 && (double)T * BN * 8.0 <= 32.0 * (1ull << 30))) {
    BALLOC(h->xin_e, (size_t)T * BN);
    BALLOC(h->xin_i, (size_t)T * BN);
  }
  if (d->gemm != 0 && d->gemm != 1) {
    free_bptt(h);
    return bs_internal_fail(BS_EINVAL, "bs_bptt_create: gemm must be 0 (bf16 split) or 1 (float32)");
  }
  h->Kp = (int)((N + 63) / 64 * 64);
  const size_t TBK = (size_t)T * B * h->Kp;
  if (d->gemm == 0) {
    BALLOC(h->wh, (size_t)N * h->Kp);
    BALLOC(h->wl, (size_t)N * h->Kp);
    for (int p = 0; p < 2; ++p) {
      BALLOC(h->ibh[p], TBK);
      BALLOC(h->ibl[p], TBK);
    }
  } else {
    BALLOC(h->wf, (size_t)N * h->Kp);
    for (int p = 0; p < 2; ++p) BALLOC(h->ibf[p], TBK);
  }
  BALLOC(h->lpart_blocks, kLossBlocks);

  BALLOC(h->sbar, (size_t)kTdSplit * h->W * BN);
  BALLOC(h->ay, BN);
  BALLOC(h->ae, BN);
  BALLOC(h->ai, BN);
  BALLOC(h->err, 1);
#undef BALLOC
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    free_bptt(h);
    return bs_internal_fail(BS_EHIP, "bs_bptt_create: hipStreamCreate failed");
  }
  // the K padding of the ibar operand rows stays zero (the adjoint writes
  // columns < N only)
  for (int p = 0; p < 2; ++p) {
    if (h->ibh[p] && (hipMemset(h->ibh[p], 0, 2 * TBK) != hipSuccess || hipMemset(h->ibl[p], 0, 2 * TBK) != hipSuccess)) {
      free_bptt(h);
      return bs_internal_fail(BS_EHIP, "bs_bptt_create: hipMemset of the GEMM operands failed");
    }
    if (h->ibf[p] && hipMemset(h->ibf[p], 0, 4 * TBK) != hipSuccess) {
      free_bptt(h);
      return bs_internal_fail(BS_EHIP, "bs_bptt_create: hipMemset of the GEMM operands failed");
    }
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
  BCHK(hipMemsetAsync(h->err, 0, sizeof(int), h->stream));
  FwdArgs a = fwd_args(h);
  a.mode = 0;
  const dim3 grid((unsigned)((N + kTile - 1) / kTile), (unsigned)B);
  for (int n0 = 0; n0 < T; n0 += h->W) {
    a.n0 = n0;
    a.nsteps = std::min(h->W, T - n0);
    hipLaunchKernelGGL(k_bptt_fwd, grid, dim3(kTile), 0, h->stream, a);
  }
  const double T_s = T * h->d.dt_ms * 1e-3;
  const double inv_count = h->d.loss == BS_LOSS_RATE_MSE ? 1.0 / (double)BN : 1.0 / ((double)BN * T);
  hipLaunchKernelGGL(k_bptt_loss_part, dim3(kLossBlocks), dim3(kTile), 0, h->stream, B, N, h->d.loss,
                     1.0 / T_s, inv_count, h->counts, h->lpart, target, h->grate, h->lpart_blocks);
  hipLaunchKernelGGL(k_bptt_loss_final, dim3(1), dim3(kTile), 0, h->stream, inv_count,
                     (const double*)h->lpart_blocks, h->d_loss);
  // the weight operand of the transposed delivery, once per forward
  hipLaunchKernelGGL(k_bptt_prep_w, dim3((unsigned)((h->Kp + kTile - 1) / kTile), (unsigned)N), dim3(kTile), 0,
                     h->stream, N, h->Kp, h->d.gemm, w_rec, h->wh, h->wl, h->wf);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(loss, h->d_loss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  int err = 0;
  BCHK(hipMemcpyAsync(&err, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  BCHK(hipStreamSynchronize(h->stream));
  if (err)
    return bs_internal_fail(BS_ECAPACITY, "bs_bptt_forward: more than 3072 arriving events in "
                                          "eight steps of one trial");
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
  ba.pstride = (size_t)h->W * BN;
  ba.nsplit = kTdSplit;
  ba.grate = h->grate;
  ba.target = h->target;
  ba.mscale = (float)(1.0 / ((double)BN * T));
  ba.ay = h->ay;
  ba.ae = h->ae;
  ba.ai = h->ai;
  ba.ibex = h->ibex;
  ba.ibin = h->ibin;
  ba.Kp = h->Kp;
  ba.mode = h->d.gemm;
  for (int p = 0; p < 2; ++p) {
    ba.ibh[p] = h->ibh[p];
    ba.ibl[p] = h->ibl[p];
    ba.ibf[p] = h->ibf[p];
  }
  TdArgs ta{};
  ta.w0 = h->d.gemm == 0 ? (const void*)h->wh : (const void*)h->wf;
  ta.w1 = h->wl;
  ta.out = h->sbar;
  ta.pstride = (size_t)h->W * BN;
  ta.nks = kTdSplit;
  ta.N = N;
  ta.Kp = h->Kp;
  ta.n_exc = nE;
  ta.nti0 = (nE + kTdBI - 1) / kTdBI;
  ta.nti1 = (nI + kTdBI - 1) / kTdBI;
  for (int seg = h->nck - 1; seg >= 0; --seg) {
    const int s0 = seg * C, s1 = std::min(T, s0 + C);
    fa.seg0 = s0;
    fa.n0 = s0;
    fa.nsteps = s1 - s0;
    if (h->xin_e)
      hipLaunchKernelGGL(k_bptt_replay, grid, dim3(kTile), 0, h->stream, fa);
    else
      hipLaunchKernelGGL(k_bptt_fwd, grid, dim3(kTile), 0, h->stream, fa);
    // windows of W steps from the end of the segment
    for (int w1 = s1; w1 > s0; w1 -= h->W) {
      const int w0 = std::max(s0, w1 - h->W);
      const int nsteps = w1 - w0;
      // sbar rows k: spikes at w0 + k reach their targets at w0 + k + d
      const int kmax = std::max(0, std::min(nsteps, T - d - w0));
      if (kmax < nsteps)
        for (int s = 0; s < kTdSplit; ++s)
          BCHK(hipMemsetAsync(h->sbar + (size_t)s * h->W * BN + (size_t)kmax * BN, 0,
                              sizeof(float) * (nsteps - kmax) * BN, h->stream));
      if (kmax > 0) {
        const int M = kmax * B;
        const size_t src = (size_t)(w0 + d) * B * h->Kp;  // first operand row: (step w0 + d, trial 0)
        for (int p = 0; p < 2; ++p) {
          ta.b0[p] = h->d.gemm == 0 ? (const void*)(h->ibh[p] + src) : (const void*)(h->ibf[p] + src);
          ta.b1[p] = h->d.gemm == 0 ? (const void*)(h->ibl[p] + src) : nullptr;
        }
        ta.M = M;
        ta.ntm = (M + kTdBM - 1) / kTdBM;
        const int tiles = (ta.nti0 + ta.nti1) * ta.ntm * ta.nks;
        const unsigned tgrid = (unsigned)((tiles + 7) / 8 * 8);
        if (h->d.gemm == 0)
          hipLaunchKernelGGL(k_bptt_tdeliver<0>, dim3(tgrid), dim3(256), 0, h->stream, ta);
        else
          hipLaunchKernelGGL(k_bptt_tdeliver<1>, dim3(tgrid), dim3(256), 0, h->stream, ta);
      }
      ba.n0 = w0;
      ba.nsteps = nsteps;
      ba.seg0 = s0;
      hipLaunchKernelGGL(k_bptt_bwd, grid, dim3(kTile), 0, h->stream, ba);
    }
  }
  WgArgs wg{};
  wg.N = N;
  wg.B = B;
  wg.T = T;
  wg.n_exc = nE;
  wg.bmT = h->bmT;
  wg.cntT = h->cntT;
  wg.ibex = h->ibex;
  wg.ibin = h->ibin;
  if (T > d) {
    wg.kind = 0;
    wg.d = d;
    wg.g = g_rec;
    hipLaunchKernelGGL(k_bptt_wgrad, dim3((unsigned)N), dim3(kTile), 0, h->stream, wg);
  } else {
    BCHK(hipMemsetAsync(g_rec, 0, sizeof(float) * (size_t)N * N, h->stream));
  }
  if (h->d.n_in > 0) {
    if (T > h->d.delay_in_steps) {
      const size_t TB = (size_t)T * B;
      hipLaunchKernelGGL(k_bptt_transpose_counts, dim3((unsigned)((TB * h->d.n_in + kTile - 1) / kTile)),
                         dim3(kTile), 0, h->stream, h->d.n_in, TB, h->cnt, h->cntT);
      wg.kind = 1;
      wg.d = h->d.delay_in_steps;
      wg.g = g_in;
      hipLaunchKernelGGL(k_bptt_wgrad, dim3((unsigned)h->d.n_in), dim3(kTile), 0, h->stream, wg);
    } else {
      BCHK(hipMemsetAsync(g_in, 0, sizeof(float) * (size_t)h->d.n_in * N, h->stream));
    }
  }
  BCHK(hipGetLastError());
  int err = 0;
  BCHK(hipMemcpyAsync(&err, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  BCHK(hipStreamSynchronize(h->stream));
  if (err) return bs_internal_fail(BS_ECAPACITY, "bs_bptt_backward: event capacity exceeded");
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
