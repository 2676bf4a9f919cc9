// bs_net.hip -- MI355X (gfx950) network-step engine behind the C ABI declared
// in include/brainspike_hip.h.
//
// Execution model (one HIP stream per network handle):
//   * Steps are grouped into windows of W = min(d_min + 1, 32) steps, d_min the
//     smallest neuron-to-neuron delay in steps. A spike emitted at step s lands
//     at step s + 1 + d (the Simulator's one-step holder lag plus the delay,
//     simulator.py:1727-1730 + delay.py:114-128), so no spike emitted inside a
//     window can reach a step of the same window.
//   * k_update advances every neuron through the W steps of a window with its
//     state in registers: it drains the arrival ring, draws the Poisson drive
//     from a counter-based stream, applies the model step, appends spikes to a
//     spike list and writes voltmeter taps.
//   * k_deliver walks the CSR row of every spike of the window and adds one
//     event per edge into the arrival-ring word (integer atomics: exact and
//     order independent, so every run is bit-reproducible). Homogeneous
//     weights are applied at drain time through S[k], the sequential sum of k
//     copies of the weight, which equals the reference's CSR event-matmul sum.
//   * Plastic (stdp_synapse) projections run at arrival time with W = 1:
//     k_stdp_post applies facilitation on the incoming edges of neurons that
//     fired in the previous step, k_stdp_pre applies depression to arriving
//     edges and deposits their weights as exact 2^-F fixed-point integers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/brainspike_hip.h"
#include "bs_math.h"
#include "bs_models.h"

using namespace bs;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(BS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

// ------------------------------------------------------------ device types
enum { FLAG_OVERFLOW = 1, FLAG_UNSTABLE = 2, FLAG_COUNT = 4 };

struct DevPop {
  int model, n, gid0, ch_begin, ch_count;
  int n_cw;          // packed count words per neuron (two 16-bit counters each)
  int n_fx;          // fixed-point channels
  int n_state;       // state variables used by the model
  long long cw_base; // first count word of this population inside a ring slot
  long long fx_base; // first fixed-point word inside a ring64 slot
  double P[BS_NPARAM];
};

struct DevChan {
  int kind, receptor;
  int cw, half;  // COUNT: word (0..n_cw) and half of the population's words
  int max_count; // COUNT: table holds S[0..max_count]
  int fx, shift; // FIXED: index and F
  long long table_off;
  double weight, L;
  int pieces, delay;
  long long tmin, tmax;
  unsigned seed;
  const int* gen_index;
};

struct Tile {
  int pop, start, count, pad;
};

struct UpdArgs {
  const Tile* tiles;
  const DevPop* pops;
  const DevChan* chans;
  const void* tables;
  void* state;  // [BS_NSTATE][Ntot]
  int* ref;
  long long Ntot;
  uint32_t* ring;
  long long CW;
  long long* ring64;
  long long FX;
  int L;
  int n0, W;
  int2* rec;
  unsigned long long* rec_count;
  long long rec_cap;
  int* flags;
  const int* vm_col;
  int n_vm;
  double* vm_buf;
  long long vm_step0;
  unsigned char* spk_hist;  // [L][Ntot] last-L-steps spike flags (plastic nets)
};

// ------------------------------------------------------------ update kernel
template <typename T>
__global__ __launch_bounds__(256) void k_update(UpdArgs a) {
  const Tile tl = a.tiles[blockIdx.x];
  const DevPop& pop = a.pops[tl.pop];
  const int model = pop.model;
  const int j = tl.start + (int)threadIdx.x;
  const bool active = (int)threadIdx.x < tl.count;
  const long long gid = (long long)pop.gid0 + j;
  const int n = pop.n;
  T* st = (T*)a.state;
  const T* tables = (const T*)a.tables;

  T s[BS_NSTATE];
  int r = 0;
  const int ns = pop.n_state;
#pragma unroll
  for (int k = 0; k < BS_NSTATE; ++k) s[k] = (T)0;
  int vmc = -1;
  if (active) {
    for (int k = 0; k < ns; ++k) s[k] = st[(long long)k * a.Ntot + gid];
    r = a.ref[gid];
    if (a.n_vm > 0) vmc = a.vm_col[gid];
  }
  const int lane = __lane_id();

  for (int w = 0; w < a.W; ++w) {
    const int step = a.n0 + w;
    const int slot = step % a.L;
    T acc0 = (T)0, acc1 = (T)0;
    if (active) {
      uint32_t words[8];
      const long long wbase = (long long)slot * a.CW + pop.cw_base + j;
      for (int q = 0; q < pop.n_cw; ++q) {
        words[q] = a.ring[wbase + (long long)q * n];
        a.ring[wbase + (long long)q * n] = 0u;
      }
      for (int c = 0; c < pop.ch_count; ++c) {
        const DevChan& ch = a.chans[pop.ch_begin + c];
        T val = (T)0;
        if (ch.kind == BS_CH_COUNT) {
          uint32_t k = (words[ch.cw] >> (16 * ch.half)) & 0xFFFFu;
          if (k > (uint32_t)ch.max_count) {  // cannot happen for a consistent ring
            atomicOr(a.flags, FLAG_COUNT);
            k = (uint32_t)ch.max_count;
          }
          val = tables[ch.table_off + k];
        } else if (ch.kind == BS_CH_POISSON) {
          const int e = ch.gen_index ? ch.gen_index[j] : j;
          const long long sg = (long long)step - 1 - ch.delay;
          uint32_t cnt = 0;
          if (e >= 0 && sg >= 0 && sg > ch.tmin && sg <= ch.tmax)
            cnt = poisson_knuth(ch.seed, (uint32_t)sg, (uint32_t)e, ch.L, ch.pieces);
          val = (T)cnt * (T)ch.weight;
        } else {  // BS_CH_FIXED
          const long long idx = (long long)slot * a.FX + pop.fx_base + (long long)ch.fx * n + j;
          const long long q = a.ring64[idx];
          a.ring64[idx] = 0;
          val = (T)__builtin_ldexp((double)q, -ch.shift);
        }
        if (ch.receptor == 0)
          acc0 = acc0 + val;
        else
          acc1 = acc1 + val;
      }
    }
    bool spike = false;
    if (active) {
      if (model == BS_IAF_PSC_DELTA) {
        spike = step_delta<T>(pop.P, s, r, acc0);
      } else if (model == BS_IAF_PSC_EXP) {
        spike = step_exp<T>(pop.P, s, r, acc0);
      } else if (model == BS_IAF_PSC_ALPHA) {
        spike = step_alpha<T>(pop.P, s, r, acc0);
      } else if (model == BS_AEIF_COND_ALPHA) {
        if constexpr (std::is_same<T, double>::value) {
          int unstable = 0;
          spike = step_aeif(pop.P, s, r, acc0, acc1, &unstable);
          if (unstable) atomicOr(a.flags, FLAG_UNSTABLE);
        }
      }
      if (vmc >= 0) a.vm_buf[(long long)(step - a.vm_step0) * a.n_vm + vmc] = (double)s[0];
      if (a.spk_hist) a.spk_hist[(long long)slot * a.Ntot + gid] = spike ? 1 : 0;
    }
    // wave-aggregated append of (gid, step)
    const unsigned long long mask = __ballot(spike);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(a.rec_count, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (spike) {
        const unsigned long long idx = base + __popcll(mask & ((1ull << lane) - 1ull));
        if ((long long)idx < a.rec_cap)
          a.rec[idx] = make_int2((int)gid, step);
        else
          atomicOr(a.flags, FLAG_OVERFLOW);
      }
    }
  }
  if (active) {
    for (int k = 0; k < ns; ++k) st[(long long)k * a.Ntot + gid] = s[k];
    a.ref[gid] = r;
  }
}

// ---------------------------------------------------------- delivery kernel
// Edge word: bits 0..23 count word inside a ring slot, bit 24 counter half,
// bits 25..31 delay in steps.
__global__ __launch_bounds__(256) void k_deliver(const int2* __restrict__ rec,
                                                 const unsigned long long* __restrict__ rec_count,
                                                 long long rec_cap, long long* win_begin,
                                                 unsigned int* done,
                                                 const long long* __restrict__ rowptr,
                                                 const uint32_t* __restrict__ edges,
                                                 uint32_t* ring, int L, long long CW) {
  const long long begin = *win_begin;
  long long end = (long long)*rec_count;
  if (end > rec_cap) end = rec_cap;
  for (long long i = begin + blockIdx.x; i < end; i += gridDim.x) {
    const int2 sp = rec[i];
    const long long e0 = rowptr[sp.x], e1 = rowptr[sp.x + 1];
    for (long long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const uint32_t wd = edges[e];
      const uint32_t cw = wd & 0xFFFFFFu;
      const uint32_t half = (wd >> 24) & 1u;
      const int d = (int)(wd >> 25);
      const int slot = (sp.y + 1 + d) % L;
      atomicAdd(&ring[(long long)slot * CW + cw], 1u << (16u * half));
    }
  }
  // the last block to finish publishes the next window's first spike index
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int prev = atomicAdd(done, 1u);
    if (prev == gridDim.x - 1) {
      *win_begin = end;
      *done = 0u;
    }
  }
}

// Heterogeneous static weights: the same walk, adding the exact fixed-point
// image q_e = rint(w_e * 2^F) into 64-bit ring words.
__global__ __launch_bounds__(256) void k_deliver_fixed(
    const int2* __restrict__ rec, const unsigned long long* __restrict__ rec_count,
    long long rec_cap, const long long* win_begin, const long long* __restrict__ rowptr,
    const uint32_t* __restrict__ tgt, const int* __restrict__ dly, const long long* __restrict__ q,
    long long* ring64, int L, long long FX) {
  const long long begin = *win_begin;
  long long end = (long long)*rec_count;
  if (end > rec_cap) end = rec_cap;
  for (long long i = begin + blockIdx.x; i < end; i += gridDim.x) {
    const int2 sp = rec[i];
    const long long e0 = rowptr[sp.x], e1 = rowptr[sp.x + 1];
    for (long long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const int slot = (sp.y + 1 + dly[e]) % L;
      atomicAdd((unsigned long long*)&ring64[(long long)slot * FX + tgt[e]],
                (unsigned long long)q[e]);
    }
  }
}


// ============================================================== host side
struct bs_net {
  int dtype = BS_F64;
  size_t tsize = 8;
  double dt = 0.1;
  hipStream_t stream = nullptr;
  int device = 0;
  // layout
  long long Ntot = 0, CW = 0, FX = 0;
  int L = 1, W = 1, d_min = 0, d_max = 0;
  std::vector<DevPop> pops;
  std::vector<DevChan> chans;
  std::vector<Tile> tiles;
  std::vector<int> pop_record;
  // device buffers
  Tile* d_tiles = nullptr;
  DevPop* d_pops = nullptr;
  DevChan* d_chans = nullptr;
  void* d_tables = nullptr;
  void* d_state = nullptr;
  int* d_ref = nullptr;
  uint32_t* d_ring = nullptr;
  long long* d_ring64 = nullptr;
  int2* d_rec = nullptr;
  unsigned long long* d_rec_count = nullptr;
  long long rec_cap = 0;
  long long* d_win_begin = nullptr;
  unsigned int* d_done = nullptr;
  int* d_flags = nullptr;
  int* d_vm_col = nullptr;
  int n_vm_dev = 0;
  double* d_vm = nullptr;
  long long vm_cap_steps = 0;
  long long* d_rowptr = nullptr;
  uint32_t* d_edges = nullptr;
  long long n_edges = 0;
  long long* d_frowptr = nullptr;
  uint32_t* d_ftgt = nullptr;
  int* d_fdly = nullptr;
  long long* d_fq = nullptr;
  long long n_fedges = 0;
  std::vector<int*> d_gen;
  // host copies
  std::vector<double> state0;  // [BS_NSTATE][Ntot]
  std::vector<int> ref0;
  std::vector<long long> rowptr_h, frowptr_h;
  std::vector<int> vm_tap_col;  // per tap (in caller order) -> device column
  long long step = 0;
  std::vector<int32_t> out_gid, out_step;
  // timing / accounting of the last run
  bool timing = false;
  double t_deliver = 0, t_update = 0, t_total = 0;
  long long n_deliver = 0, n_update = 0;
  double b_deliver = 0, b_update = 0;
  long long n_events = 0;
};

// Public entry points keep the C linkage of their declarations in the header.
int bs_abi_version(void) { return BS_ABI_VERSION; }
const char* bs_last_error(void) { return g_err.c_str(); }
int bs_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static void free_net(bs_net* net) {
  if (!net) return;
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(net->d_tiles); F(net->d_pops); F(net->d_chans); F(net->d_tables); F(net->d_state);
  F(net->d_ref); F(net->d_ring); F(net->d_ring64); F(net->d_rec); F(net->d_rec_count);
  F(net->d_win_begin); F(net->d_done); F(net->d_flags); F(net->d_vm_col); F(net->d_vm);
  F(net->d_rowptr); F(net->d_edges); F(net->d_frowptr); F(net->d_ftgt); F(net->d_fdly);
  F(net->d_fq);
  for (int* p : net->d_gen) F(p);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

int bs_net_destroy(bs_net* net) {
  free_net(net);
  return BS_OK;
}

static int model_nstate(int model, const double* P) {
  switch (model) {
    case BS_IAF_PSC_DELTA: return P[D_REFIN] != 0.0 ? 2 : 1;
    case BS_IAF_PSC_EXP: return 5;
    case BS_IAF_PSC_ALPHA: return 6;
    case BS_AEIF_COND_ALPHA: return 8;
  }
  return -1;
}

template <typename X>
static int dmalloc(X** p, size_t count) {
  *p = nullptr;
  if (count == 0) return BS_OK;
  hipError_t e = hipMalloc((void**)p, count * sizeof(X));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return BS_OK;
}

#define CHK(x)                        \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != BS_OK) {               \
      free_net(net);                  \
      return rc_;                     \
    }                                 \
  } while (0)
#define HCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      free_net(net);                                                                     \
      return fail(BS_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));              \
    }                                                                                    \
  } while (0)
#define BAD(msg)                        \
  do {                                  \
    free_net(net);                      \
    return fail(BS_EINVAL, (msg));      \
  } while (0)

static int upload_state(bs_net* net);

int bs_net_create(const bs_net_desc* d, bs_net** out) {
  if (!d || !out) return fail(BS_EINVAL, "bs_net_create: null argument");
  *out = nullptr;
  if (d->abi_version != BS_ABI_VERSION)
    return fail(BS_EINVAL, "bs_net_create: ABI version mismatch");
  if (d->dtype != BS_F64 && d->dtype != BS_F32)
    return fail(BS_EINVAL, "bs_net_create: dtype must be BS_F64 or BS_F32");
  if (!(d->dt_ms > 0)) return fail(BS_EINVAL, "bs_net_create: dt must be > 0");
  if (d->n_pops < 1 || !d->pops) return fail(BS_EINVAL, "bs_net_create: no populations");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(BS_ENODEV, "bs_net_create: no HIP device visible");

  bs_net* net = new bs_net();
  net->dtype = d->dtype;
  net->tsize = d->dtype == BS_F64 ? 8 : 4;
  net->dt = d->dt_ms;
  HCHK(hipGetDevice(&net->device));
  HCHK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));

  // ---- populations and channels
  long long gid = 0, cw = 0, fx = 0;
  net->pops.resize(d->n_pops);
  net->chans.resize(d->n_channels);
  for (int c = 0; c < d->n_channels; ++c) memset(&net->chans[c], 0, sizeof(DevChan));
  std::vector<int> chan_pop(d->n_channels, -1);
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    DevPop& dp = net->pops[p];
    memset(&dp, 0, sizeof(dp));
    dp.model = sp.model;
    dp.n = sp.n;
    if (sp.n < 0) BAD("population size must be >= 0");
    dp.gid0 = (int)gid;
    memcpy(dp.P, sp.params, sizeof(dp.P));
    dp.n_state = model_nstate(sp.model, sp.params);
    if (dp.n_state < 0) BAD("unknown neuron model id " + std::to_string(sp.model));
    if (sp.model == BS_AEIF_COND_ALPHA && d->dtype != BS_F64)
      BAD("aeif_cond_alpha runs in float64 only");
    if (!sp.state0) BAD("population state0 must be provided");
    dp.ch_begin = sp.ch_begin;
    dp.ch_count = sp.ch_count;
    if (sp.ch_begin < 0 || sp.ch_count < 0 || sp.ch_begin + sp.ch_count > d->n_channels)
      BAD("population channel range out of bounds");
    int ncount = 0, nfx = 0;
    for (int c = sp.ch_begin; c < sp.ch_begin + sp.ch_count; ++c) {
      const bs_channel& ch = d->channels[c];
      DevChan& dc = net->chans[c];
      chan_pop[c] = p;
      dc.kind = ch.kind;
      dc.receptor = ch.receptor;
      if (ch.receptor < 0 || ch.receptor > 1) BAD("channel receptor must be 0 or 1");
      if (ch.receptor == 1 && sp.model != BS_AEIF_COND_ALPHA)
        BAD("receptor 1 only exists on aeif_cond_alpha");
      dc.weight = ch.weight;
      if (ch.kind == BS_CH_COUNT) {
        if (ch.max_count < 0 || ch.max_count > 65535)
          BAD("COUNT channel max_count must be in [0, 65535]");
        if (!d->table_off || d->table_off[c] < 0 ||
            d->table_off[c] + ch.max_count >= d->n_tables)
          BAD("COUNT channel table out of range");
        dc.table_off = d->table_off[c];
        dc.max_count = ch.max_count;
        dc.cw = ncount / 2;
        dc.half = ncount % 2;
        ++ncount;
      } else if (ch.kind == BS_CH_FIXED) {
        dc.fx = nfx++;
        dc.shift = ch.fixed_shift;
        if (ch.fixed_shift < 0 || ch.fixed_shift > 1000) BAD("bad fixed_shift");
      } else if (ch.kind == BS_CH_POISSON) {
        if (!(ch.lam >= 0)) BAD("Poisson lambda must be >= 0");
        int pieces = (int)std::ceil(ch.lam / 16.0);
        if (pieces < 1) pieces = 1;
        dc.pieces = pieces;
        dc.L = std::exp(-(ch.lam / (double)pieces));
        dc.tmin = ch.t_min_step;
        dc.tmax = ch.t_max_step;
        dc.seed = ch.seed;
        dc.delay = ch.delay_steps;
        if (ch.delay_steps < 0) BAD("Poisson delay must be >= 0");
        if (ch.lam == 0.0) dc.tmax = -1;  // rate <= 0: all-zero output
        if (ch.gen_index) {
          int* g = nullptr;
          CHK(dmalloc(&g, (size_t)sp.n));
          net->d_gen.push_back(g);
          HCHK(hipMemcpy(g, ch.gen_index, sizeof(int) * sp.n, hipMemcpyHostToDevice));
          dc.gen_index = g;
        }
      } else {
        BAD("unknown channel kind");
      }
    }
    dp.n_cw = (ncount + 1) / 2;
    if (dp.n_cw > 8) BAD("at most 16 COUNT channels per population");
    dp.n_fx = nfx;
    dp.cw_base = cw;
    dp.fx_base = fx;
    cw += (long long)dp.n_cw * sp.n;
    fx += (long long)nfx * sp.n;
    gid += sp.n;
    net->pop_record.push_back(sp.record_spikes);
  }
  net->Ntot = gid;
  net->CW = cw;
  net->FX = fx;
  if (gid >= (1ll << 31)) BAD("more than 2^31 neurons");
  if (cw >= (1ll << 24)) BAD("count ring words exceed 2^24 (too many neurons x channels)");

  // ---- projections -> CSR rows by pre gid
  std::vector<long long> cnt(gid + 1, 0), fcnt(gid + 1, 0);
  int dmin = 1 << 30, dmax = 0;
  bool have_plastic = false;
  for (int q = 0; q < d->n_projs; ++q) {
    const bs_projection& pj = d->projs[q];
    if (pj.pre_pop < 0 || pj.pre_pop >= d->n_pops || pj.post_pop < 0 || pj.post_pop >= d->n_pops)
      BAD("projection population index out of range");
    if (pj.channel < 0 || pj.channel >= d->n_channels || chan_pop[pj.channel] != pj.post_pop)
      BAD("projection channel does not belong to the post population");
    if (pj.plastic) have_plastic = true;
    if (pj.plastic) BAD("plastic projections are not supported by this engine build");
    const int kind = d->channels[pj.channel].kind;
    if (kind != BS_CH_COUNT && kind != BS_CH_FIXED) BAD("projection channel must be COUNT or FIXED");
    if (kind == BS_CH_FIXED && !pj.weight) BAD("FIXED projection needs per-edge weights");
    const int npre = d->pops[pj.pre_pop].n, npost = d->pops[pj.post_pop].n;
    for (long long e = 0; e < pj.n_edges; ++e) {
      const int a = pj.pre[e], b = pj.post[e];
      if (a < 0 || a >= npre || b < 0 || b >= npost) BAD("edge index out of range");
      const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
      if (dd < 0) BAD("delay must be >= 0 steps");
      if (kind == BS_CH_COUNT && dd > 127) BAD("delay above 127 steps is not supported");
      dmin = std::min(dmin, dd);
      dmax = std::max(dmax, dd);
      const long long g = net->pops[pj.pre_pop].gid0 + a;
      if (kind == BS_CH_COUNT)
        cnt[g + 1]++;
      else
        fcnt[g + 1]++;
    }
  }
  (void)have_plastic;
  if (dmin > dmax) dmin = dmax = 0;
  for (long long g = 0; g < gid; ++g) {
    cnt[g + 1] += cnt[g];
    fcnt[g + 1] += fcnt[g];
  }
  net->n_edges = cnt[gid];
  net->n_fedges = fcnt[gid];
  std::vector<uint32_t> edges(net->n_edges);
  std::vector<uint32_t> ftgt(net->n_fedges);
  std::vector<int> fdly(net->n_fedges);
  std::vector<long long> fq(net->n_fedges);
  {
    std::vector<long long> pos(cnt.begin(), cnt.end() - 1), fpos(fcnt.begin(), fcnt.end() - 1);
    for (int q = 0; q < d->n_projs; ++q) {
      const bs_projection& pj = d->projs[q];
      const bs_channel& ch = d->channels[pj.channel];
      const DevPop& post = net->pops[pj.post_pop];
      const DevChan& dc = net->chans[pj.channel];
      for (long long e = 0; e < pj.n_edges; ++e) {
        const long long g = net->pops[pj.pre_pop].gid0 + pj.pre[e];
        const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
        if (ch.kind == BS_CH_COUNT) {
          const long long w = post.cw_base + (long long)dc.cw * post.n + pj.post[e];
          edges[pos[g]++] = (uint32_t)w | ((uint32_t)dc.half << 24) | ((uint32_t)dd << 25);
        } else {
          const long long f = post.fx_base + (long long)dc.fx * post.n + pj.post[e];
          const long long k = fpos[g]++;
          ftgt[k] = (uint32_t)f;
          fdly[k] = dd;
          fq[k] = (long long)std::llrint(std::ldexp(pj.weight[e], ch.fixed_shift));
        }
      }
    }
  }
  net->d_min = dmin;
  net->d_max = dmax;
  int W = dmin + 1;
  const int maxw = d->max_window > 0 ? d->max_window : 32;
  if (W > maxw) W = maxw;
  if (W < 1) W = 1;
  net->W = W;
  net->L = dmax + W + 1;
  net->rowptr_h = cnt;
  net->frowptr_h = fcnt;

  // ---- tiles (<=256 neurons of one population each)
  for (int p = 0; p < d->n_pops; ++p)
    for (int s0 = 0; s0 < net->pops[p].n; s0 += 256)
      net->tiles.push_back(Tile{p, s0, std::min(256, net->pops[p].n - s0), 0});

  // ---- voltmeter columns
  std::vector<int> vm_col(gid, -1);
  int ncol = 0;
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    for (int k = 0; k < sp.n_vm; ++k) {
      const int li = sp.vm_idx[k];
      if (li < 0 || li >= sp.n) BAD("voltmeter index out of range");
      const long long g = net->pops[p].gid0 + li;
      if (vm_col[g] < 0) vm_col[g] = ncol++;
      net->vm_tap_col.push_back(vm_col[g]);
    }
  }
  net->n_vm_dev = ncol;

  // ---- state0 / ref0 (host copies for reset)
  net->state0.assign((size_t)BS_NSTATE * gid, 0.0);
  net->ref0.assign(gid, 0);
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    const long long g0 = net->pops[p].gid0;
    for (int k = 0; k < BS_NSTATE; ++k)
      for (int i = 0; i < sp.n; ++i) net->state0[(size_t)k * gid + g0 + i] = sp.state0[(size_t)k * sp.n + i];
  }

  // ---- device allocations
  const size_t ts = net->tsize;
  CHK(dmalloc(&net->d_tiles, net->tiles.size()));
  CHK(dmalloc(&net->d_pops, net->pops.size()));
  CHK(dmalloc(&net->d_chans, std::max<size_t>(1, net->chans.size())));
  CHK(dmalloc((char**)&net->d_tables, std::max<long long>(1, d->n_tables) * ts));
  CHK(dmalloc((char**)&net->d_state, (size_t)BS_NSTATE * gid * ts));
  CHK(dmalloc(&net->d_ref, gid));
  CHK(dmalloc(&net->d_ring, (size_t)net->L * std::max<long long>(1, cw)));
  CHK(dmalloc(&net->d_ring64, (size_t)net->L * std::max<long long>(1, fx)));
  CHK(dmalloc(&net->d_rowptr, gid + 1));
  CHK(dmalloc(&net->d_edges, std::max<long long>(1, net->n_edges)));
  CHK(dmalloc(&net->d_frowptr, gid + 1));
  CHK(dmalloc(&net->d_ftgt, std::max<long long>(1, net->n_fedges)));
  CHK(dmalloc(&net->d_fdly, std::max<long long>(1, net->n_fedges)));
  CHK(dmalloc(&net->d_fq, std::max<long long>(1, net->n_fedges)));
  CHK(dmalloc(&net->d_rec_count, 1));
  CHK(dmalloc(&net->d_win_begin, 1));
  CHK(dmalloc(&net->d_done, 1));
  CHK(dmalloc(&net->d_flags, 1));
  CHK(dmalloc(&net->d_vm_col, std::max<long long>(1, gid)));
  // spike list: worst case every neuron fires every step of K windows
  {
    const long long per_win = std::max<long long>(1, gid) * W;
    long long k = (64ll << 20) / per_win;
    if (k < 1) k = 1;
    if (k > 256) k = 256;
    net->rec_cap = per_win * k;
    CHK(dmalloc(&net->d_rec, net->rec_cap));
    net->vm_cap_steps = k * W;
    CHK(dmalloc(&net->d_vm, (size_t)std::max(1, ncol) * net->vm_cap_steps));
  }
  HCHK(hipMemcpy(net->d_tiles, net->tiles.data(), sizeof(Tile) * net->tiles.size(), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(net->d_pops, net->pops.data(), sizeof(DevPop) * net->pops.size(), hipMemcpyHostToDevice));
  if (!net->chans.empty())
    HCHK(hipMemcpy(net->d_chans, net->chans.data(), sizeof(DevChan) * net->chans.size(), hipMemcpyHostToDevice));
  if (d->n_tables > 0) {
    if (ts == 8) {
      HCHK(hipMemcpy(net->d_tables, d->tables, 8 * d->n_tables, hipMemcpyHostToDevice));
    } else {
      std::vector<float> tf(d->n_tables);
      for (long long i = 0; i < d->n_tables; ++i) tf[i] = (float)d->tables[i];
      HCHK(hipMemcpy(net->d_tables, tf.data(), 4 * d->n_tables, hipMemcpyHostToDevice));
    }
  }
  HCHK(hipMemcpy(net->d_rowptr, cnt.data(), sizeof(long long) * (gid + 1), hipMemcpyHostToDevice));
  if (net->n_edges)
    HCHK(hipMemcpy(net->d_edges, edges.data(), 4 * net->n_edges, hipMemcpyHostToDevice));
  HCHK(hipMemcpy(net->d_frowptr, fcnt.data(), sizeof(long long) * (gid + 1), hipMemcpyHostToDevice));
  if (net->n_fedges) {
    HCHK(hipMemcpy(net->d_ftgt, ftgt.data(), 4 * net->n_fedges, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(net->d_fdly, fdly.data(), 4 * net->n_fedges, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(net->d_fq, fq.data(), 8 * net->n_fedges, hipMemcpyHostToDevice));
  }
  if (gid) HCHK(hipMemcpy(net->d_vm_col, vm_col.data(), 4 * gid, hipMemcpyHostToDevice));
  {
    int rc = upload_state(net);
    if (rc != BS_OK) {
      free_net(net);
      return rc;
    }
  }
  *out = net;
  return BS_OK;
}

static int upload_state(bs_net* net) {
  const long long gid = net->Ntot;
  if (net->tsize == 8) {
    HIPCHK(hipMemcpy(net->d_state, net->state0.data(), 8 * net->state0.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(net->state0.size());
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)net->state0[i];
    HIPCHK(hipMemcpy(net->d_state, f.data(), 4 * f.size(), hipMemcpyHostToDevice));
  }
  if (gid) HIPCHK(hipMemcpy(net->d_ref, net->ref0.data(), 4 * gid, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(net->d_ring, 0, 4 * (size_t)net->L * std::max<long long>(1, net->CW)));
  HIPCHK(hipMemset(net->d_ring64, 0, 8 * (size_t)net->L * std::max<long long>(1, net->FX)));
  HIPCHK(hipMemset(net->d_done, 0, sizeof(unsigned int)));
  HIPCHK(hipMemset(net->d_flags, 0, sizeof(int)));
  HIPCHK(hipDeviceSynchronize());
  net->step = 0;
  return BS_OK;
}

int bs_net_reset(bs_net* net) {
  if (!net) return fail(BS_EINVAL, "bs_net_reset: null handle");
  net->out_gid.clear();
  net->out_step.clear();
  return upload_state(net);
}

int bs_net_step(bs_net* net, int64_t* step) {
  if (!net || !step) return fail(BS_EINVAL, "bs_net_step: null argument");
  *step = net->step;
  return BS_OK;
}

template <typename T>
static void launch_update(bs_net* net, int n0, int W, long long vm_step0) {
  UpdArgs a;
  a.tiles = net->d_tiles;
  a.pops = net->d_pops;
  a.chans = net->d_chans;
  a.tables = net->d_tables;
  a.state = net->d_state;
  a.ref = net->d_ref;
  a.Ntot = net->Ntot;
  a.ring = net->d_ring;
  a.CW = net->CW;
  a.ring64 = net->d_ring64;
  a.FX = net->FX;
  a.L = net->L;
  a.n0 = n0;
  a.W = W;
  a.rec = net->d_rec;
  a.rec_count = net->d_rec_count;
  a.rec_cap = net->rec_cap;
  a.flags = net->d_flags;
  a.vm_col = net->d_vm_col;
  a.n_vm = net->n_vm_dev;
  a.vm_buf = net->d_vm;
  a.vm_step0 = vm_step0;
  a.spk_hist = nullptr;
  hipLaunchKernelGGL(k_update<T>, dim3((unsigned)net->tiles.size()), dim3(256), 0, net->stream, a);
}

static const unsigned kDeliverBlocks = 1024;

static void launch_deliver(bs_net* net) {
  if (net->n_fedges)
    hipLaunchKernelGGL(k_deliver_fixed, dim3(kDeliverBlocks), dim3(256), 0, net->stream, net->d_rec,
                       net->d_rec_count, net->rec_cap, net->d_win_begin, net->d_frowptr, net->d_ftgt,
                       net->d_fdly, net->d_fq, net->d_ring64, net->L, net->FX);
  hipLaunchKernelGGL(k_deliver, dim3(kDeliverBlocks), dim3(256), 0, net->stream, net->d_rec,
                     net->d_rec_count, net->rec_cap, net->d_win_begin, net->d_done, net->d_rowptr,
                     net->d_edges, net->d_ring, net->L, net->CW);
}

int bs_net_set_timing(bs_net* net, int on) {
  if (!net) return fail(BS_EINVAL, "bs_net_set_timing: null handle");
  net->timing = on != 0;
  return BS_OK;
}

int bs_net_run(bs_net* net, int64_t n_steps, double* vm_out) {
  if (!net) return fail(BS_EINVAL, "bs_net_run: null handle");
  if (n_steps < 0) return fail(BS_EINVAL, "bs_net_run: n_steps must be >= 0");
  if (net->step + n_steps >= (1ll << 31)) return fail(BS_EINVAL, "bs_net_run: step counter overflow");
  HIPCHK(hipSetDevice(net->device));
  net->out_gid.clear();
  net->out_step.clear();
  net->t_deliver = net->t_update = net->t_total = 0;
  net->n_deliver = net->n_update = 0;
  net->b_deliver = net->b_update = 0;
  net->n_events = 0;
  const int ntaps = (int)net->vm_tap_col.size();
  const long long chunk_max = net->vm_cap_steps;  // = K * W
  std::vector<int2> hrec;
  std::vector<double> hvm;
  hipEvent_t ev_all0, ev_all1;
  HIPCHK(hipEventCreate(&ev_all0));
  HIPCHK(hipEventCreate(&ev_all1));
  std::vector<hipEvent_t> evs;
  auto t_wall0 = std::chrono::steady_clock::now();
  HIPCHK(hipEventRecord(ev_all0, net->stream));
  long long done = 0;
  int rc = BS_OK;
  while (done < n_steps) {
    const long long chunk = std::min<long long>(chunk_max, n_steps - done);
    const long long c0 = net->step;
    HIPCHK(hipMemsetAsync(net->d_rec_count, 0, sizeof(unsigned long long), net->stream));
    HIPCHK(hipMemsetAsync(net->d_win_begin, 0, sizeof(long long), net->stream));
    const int nwin = (int)((chunk + net->W - 1) / net->W);
    if (net->timing) {
      for (size_t i = evs.size(); i < (size_t)nwin * 3; ++i) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        evs.push_back(e);
      }
    }
    for (int w = 0; w < nwin; ++w) {
      const long long n0 = c0 + (long long)w * net->W;
      const int Wc = (int)std::min<long long>(net->W, c0 + chunk - n0);
      if (net->timing) HIPCHK(hipEventRecord(evs[3 * w], net->stream));
      if (net->tsize == 8)
        launch_update<double>(net, (int)n0, Wc, c0);
      else
        launch_update<float>(net, (int)n0, Wc, c0);
      if (net->timing) HIPCHK(hipEventRecord(evs[3 * w + 1], net->stream));
      launch_deliver(net);
      if (net->timing) HIPCHK(hipEventRecord(evs[3 * w + 2], net->stream));
      HIPCHK(hipGetLastError());
      net->n_update++;
      net->n_deliver++;
      // algorithmic update bytes: state in+out, ref in+out, ring words read+clear
      for (const DevPop& p : net->pops) {
        net->b_update += (double)p.n * (2.0 * p.n_state * net->tsize + 8.0 +
                                        (double)Wc * (8.0 * p.n_cw + 16.0 * p.n_fx));
      }
    }
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, net->d_rec_count, sizeof(cnt), hipMemcpyDeviceToHost, net->stream));
    int flags = 0;
    HIPCHK(hipMemcpyAsync(&flags, net->d_flags, sizeof(int), hipMemcpyDeviceToHost, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    if (flags & FLAG_OVERFLOW) {
      rc = fail(BS_ECAPACITY, "spike list overflow");
      break;
    }
    if (flags & FLAG_COUNT) {
      rc = fail(BS_ECAPACITY, "arrival-ring event count exceeded the channel's in-degree");
      break;
    }
    if (flags & FLAG_UNSTABLE) {
      rc = fail(BS_ENUMERIC, "Numerical instability in aeif_cond_alpha dynamics.");
      break;
    }
    if (net->timing) {
      for (int w = 0; w < nwin; ++w) {
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, evs[3 * w], evs[3 * w + 1]));
        HIPCHK(hipEventElapsedTime(&b, evs[3 * w + 1], evs[3 * w + 2]));
        net->t_update += a;
        net->t_deliver += b;
      }
    }
    hrec.resize(cnt);
    if (cnt) HIPCHK(hipMemcpy(hrec.data(), net->d_rec, sizeof(int2) * cnt, hipMemcpyDeviceToHost));
    for (unsigned long long i = 0; i < cnt; ++i) {
      const int g = hrec[i].x;
      const long long row = net->rowptr_h[g + 1] - net->rowptr_h[g];
      const long long frow = net->frowptr_h[g + 1] - net->frowptr_h[g];
      net->n_events += row + frow;
      net->b_deliver += 8.0 + 16.0 + 4.0 * row + (frow ? 16.0 + 16.0 * frow : 0.0);
      // population of g (few populations: linear scan)
      int p = 0;
      while (p + 1 < (int)net->pops.size() && g >= net->pops[p + 1].gid0) ++p;
      if (net->pop_record[p]) {
        net->out_gid.push_back(g);
        net->out_step.push_back(hrec[i].y);
      }
    }
    net->b_update += 8.0 * cnt;
    if (vm_out && ntaps > 0) {
      hvm.resize((size_t)chunk * std::max(1, net->n_vm_dev));
      HIPCHK(hipMemcpy(hvm.data(), net->d_vm, 8 * hvm.size(), hipMemcpyDeviceToHost));
      for (long long t = 0; t < chunk; ++t)
        for (int k = 0; k < ntaps; ++k)
          vm_out[(done + t) * ntaps + k] = hvm[(size_t)t * net->n_vm_dev + net->vm_tap_col[k]];
    }
    net->step += chunk;
    done += chunk;
  }
  HIPCHK(hipEventRecord(ev_all1, net->stream));
  HIPCHK(hipEventSynchronize(ev_all1));
  float tot = 0;
  HIPCHK(hipEventElapsedTime(&tot, ev_all0, ev_all1));
  (void)t_wall0;
  net->t_total = tot;
  for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ev_all0);
  (void)hipEventDestroy(ev_all1);
  // sort recorded spikes by (step, gid)
  {
    std::vector<long long> key(net->out_gid.size());
    for (size_t i = 0; i < key.size(); ++i) key[i] = ((long long)net->out_step[i] << 32) | (uint32_t)net->out_gid[i];
    std::sort(key.begin(), key.end());
    for (size_t i = 0; i < key.size(); ++i) {
      net->out_step[i] = (int32_t)(key[i] >> 32);
      net->out_gid[i] = (int32_t)(key[i] & 0xFFFFFFFF);
    }
  }
  return rc;
}

int bs_net_spike_count(bs_net* net, int64_t* n) {
  if (!net || !n) return fail(BS_EINVAL, "bs_net_spike_count: null argument");
  *n = (int64_t)net->out_gid.size();
  return BS_OK;
}

int bs_net_copy_spikes(bs_net* net, int32_t* gid, int32_t* step, int64_t cap) {
  if (!net || !gid || !step) return fail(BS_EINVAL, "bs_net_copy_spikes: null argument");
  if (cap < (int64_t)net->out_gid.size()) return fail(BS_ECAPACITY, "bs_net_copy_spikes: cap too small");
  memcpy(gid, net->out_gid.data(), 4 * net->out_gid.size());
  memcpy(step, net->out_step.data(), 4 * net->out_step.size());
  return BS_OK;
}

int bs_net_get_state(bs_net* net, int32_t pop, int32_t var, double* out) {
  if (!net || !out) return fail(BS_EINVAL, "bs_net_get_state: null argument");
  if (pop < 0 || pop >= (int)net->pops.size()) return fail(BS_EINVAL, "bs_net_get_state: bad population");
  if (var < -1 || var >= BS_NSTATE) return fail(BS_EINVAL, "bs_net_get_state: bad variable");
  const DevPop& p = net->pops[pop];
  HIPCHK(hipStreamSynchronize(net->stream));
  if (var == -1) {  // refractory counter
    std::vector<int> r(p.n);
    if (p.n) HIPCHK(hipMemcpy(r.data(), net->d_ref + p.gid0, 4 * (size_t)p.n, hipMemcpyDeviceToHost));
    for (int i = 0; i < p.n; ++i) out[i] = r[i];
    return BS_OK;
  }
  const size_t off = (size_t)var * net->Ntot + p.gid0;
  if (net->tsize == 8) {
    if (p.n) HIPCHK(hipMemcpy(out, (double*)net->d_state + off, 8 * (size_t)p.n, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(p.n);
    if (p.n) HIPCHK(hipMemcpy(f.data(), (float*)net->d_state + off, 4 * (size_t)p.n, hipMemcpyDeviceToHost));
    for (int i = 0; i < p.n; ++i) out[i] = f[i];
  }
  return BS_OK;
}

int bs_net_get_weights(bs_net* net, int32_t proj, double* out) {
  (void)proj;
  (void)out;
  if (!net) return fail(BS_EINVAL, "bs_net_get_weights: null handle");
  return fail(BS_EINVAL, "bs_net_get_weights: projection is not plastic");
}

int bs_net_last_timing(bs_net* net, double* deliver_ms, double* update_ms, double* total_ms,
                       int64_t* n_deliver, int64_t* n_update) {
  if (!net) return fail(BS_EINVAL, "bs_net_last_timing: null handle");
  if (deliver_ms) *deliver_ms = net->t_deliver;
  if (update_ms) *update_ms = net->t_update;
  if (total_ms) *total_ms = net->t_total;
  if (n_deliver) *n_deliver = net->n_deliver;
  if (n_update) *n_update = net->n_update;
  return BS_OK;
}

int bs_net_last_bytes(bs_net* net, double* deliver_bytes, double* update_bytes, int64_t* n_events) {
  if (!net) return fail(BS_EINVAL, "bs_net_last_bytes: null handle");
  if (deliver_bytes) *deliver_bytes = net->b_deliver;
  if (update_bytes) *update_bytes = net->b_update;
  if (n_events) *n_events = net->n_events;
  return BS_OK;
}

// ------------------------------------------------------------- test hooks
__global__ void k_dbg_math(int op, long long n, const double* a, const double* b, double* o) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 0) o[i] = exp_p(a[i]);
  else if (op == 1) o[i] = log_p(a[i]);
  else if (op == 2) o[i] = pow_p(a[i], b[i]);
}
__global__ void k_dbg_rng(long long n, const uint32_t* k0, const uint32_t* k1, const uint32_t* c0,
                          const uint32_t* c1, uint32_t* o0, uint32_t* o1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  threefry2x32(k0[i], k1[i], c0[i], c1[i], o0[i], o1[i]);
}
__global__ void k_dbg_poisson(long long n, const uint32_t* seed, const uint32_t* step,
                              const uint32_t* elem, double L, int pieces, uint32_t* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = poisson_knuth(seed[i], step[i], elem[i], L, pieces);
}

template <typename X>
static int to_dev(X** d, const X* h, long long n) {
  if (n <= 0) {
    *d = nullptr;
    return BS_OK;
  }
  int rc = dmalloc(d, (size_t)n);
  if (rc) return rc;
  if (h) HIPCHK(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
  return BS_OK;
}

int bs_debug_math(int32_t op, int64_t n, const double* a, const double* b, double* out) {
  if (op < 0 || op > 2 || n < 0 || !a || !out) return fail(BS_EINVAL, "bs_debug_math: bad args");
  if (n == 0) return BS_OK;
  double *da, *db = nullptr, *dout;
  int rc = to_dev(&da, a, n);
  if (rc) return rc;
  if (b) to_dev(&db, b, n);
  else to_dev(&db, a, n);
  to_dev(&dout, (const double*)nullptr, n);
  hipLaunchKernelGGL(k_dbg_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op, (long long)n, da, db, dout);
  hipError_t e = hipMemcpy(out, dout, 8 * n, hipMemcpyDeviceToHost);
  (void)hipFree(da); (void)hipFree(db); (void)hipFree(dout);
  if (e != hipSuccess) return fail(BS_EHIP, hipGetErrorString(e));
  return BS_OK;
}

int bs_debug_threefry(int64_t n, const uint32_t* k0, const uint32_t* k1, const uint32_t* c0,
                      const uint32_t* c1, uint32_t* o0, uint32_t* o1) {
  if (n < 0 || !k0 || !k1 || !c0 || !c1 || !o0 || !o1) return fail(BS_EINVAL, "bs_debug_threefry: bad args");
  if (n == 0) return BS_OK;
  uint32_t *a, *b, *c, *dd, *x, *y;
  to_dev(&a, k0, n); to_dev(&b, k1, n); to_dev(&c, c0, n); to_dev(&dd, c1, n);
  to_dev(&x, (const uint32_t*)nullptr, n); to_dev(&y, (const uint32_t*)nullptr, n);
  hipLaunchKernelGGL(k_dbg_rng, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (long long)n, a, b, c, dd, x, y);
  hipError_t e1 = hipMemcpy(o0, x, 4 * n, hipMemcpyDeviceToHost);
  hipError_t e2 = hipMemcpy(o1, y, 4 * n, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); (void)hipFree(dd); (void)hipFree(x); (void)hipFree(y);
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(BS_EHIP, "bs_debug_threefry: copy failed");
  return BS_OK;
}

int bs_debug_poisson(int64_t n, const uint32_t* seed, const uint32_t* step, const uint32_t* elem,
                     double lam, uint32_t* out) {
  if (n < 0 || !seed || !step || !elem || !out || !(lam >= 0)) return fail(BS_EINVAL, "bs_debug_poisson: bad args");
  if (n == 0) return BS_OK;
  int pieces = (int)std::ceil(lam / 16.0);
  if (pieces < 1) pieces = 1;
  const double L = std::exp(-(lam / (double)pieces));
  uint32_t *a, *b, *c, *o;
  to_dev(&a, seed, n); to_dev(&b, step, n); to_dev(&c, elem, n); to_dev(&o, (const uint32_t*)nullptr, n);
  hipLaunchKernelGGL(k_dbg_poisson, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (long long)n, a, b, c, L, pieces, o);
  hipError_t e = hipMemcpy(out, o, 4 * n, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); (void)hipFree(o);
  if (e != hipSuccess) return fail(BS_EHIP, hipGetErrorString(e));
  return BS_OK;
}

