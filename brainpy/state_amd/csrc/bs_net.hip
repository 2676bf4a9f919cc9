// bs_net.hip -- MI355X (gfx950) network-step engine behind the C ABI declared
// in include/brainspike_hip.h.
//
// Execution model (one HIP stream per network handle):
//   * Steps are grouped into windows of W = min(d_min + 1, 32) steps, d_min the
//     smallest neuron-to-neuron delay in steps. A spike emitted at step s lands
//     at step s + 1 + d (the Simulator's one-step holder lag plus the delay,
//     simulator.py:1727-1730 + delay.py:114-128), so no spike emitted inside a
//     window can reach a step of the same window.
//   * One launch of k_window advances one window. Each block owns a tile of
//     TS post neurons of one population (XCD-aware tile order) and runs five
//     phases: stage (parameters, channels, count tables, Poisson CDFs into
//     LDS); gather (every spike whose delayed arrival falls in the window
//     walks its (pre, tile) edge segment and adds one event per edge to an
//     LDS counter per (step, channel, lane), integer atomics, so the result is
//     exact and order independent); inputs (per step and lane, the channels in
//     the reference's summation order: S[k] tables for homogeneous weights,
//     exact 2^-F fixed-point sums for per-edge weights, count * weight for
//     Poisson drive drawn from a counter-based Threefry stream; dc_generator
//     currents summed apart, for the models' current-input seam); update (the
//     model step for W steps with the state in registers); emit (per-step spike
//     buckets for the following windows, the spike record, voltmeter taps).
//   * When the tiles leave CUs idle, the launch carries extra producer blocks
//     that draw the next window's Poisson counts (the first Poisson channel
//     of each population) into a two-window ring; the tiles then read counts
//     in the stage phase instead of drawing them in the inputs phase.
//   * Full chunks of windows are captured once into a HIP graph and replayed;
//     the kernels read the chunk's first step from device memory. The spike
//     records of a chunk go to one of two buffers: while the device runs
//     chunk c, a second stream sorts chunk c - 1's records and the host
//     decodes them from pinned memory.
//   * Plastic (stdp_synapse) projections run inside the same window kernel.
//     Every weight update of an edge is caused by its post neuron (a post
//     spike facilitates) or by its own arrival at the post neuron (depression
//     and delivery), so the thread of the post neuron owns its incoming edges
//     for the window: the arrivals of the window are gathered into LDS lists
//     per (step, lane), and the update loop applies, per arriving edge, the
//     facilitations still pending from the neuron's earlier spikes in the
//     window, then the arrival's facilitate/depress, and deposits the new
//     weight as an exact 2^-F fixed-point sum into that step's input. After
//     the loop the block facilitates the remaining incoming edges of the
//     neurons that fired (those with no later arrival in the window). The pre
//     traces are a per-neuron history ring written by the pre neuron's thread;
//     everything read from it lies before the window (W <= d_min + 1).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../../include/brainspike_hip.h"
#include "bs_math.h"
#include "bs_models.h"

using namespace bs;

// Gather batch: spikes per thread per pass, and 16-byte edge chunks per spike
// and round (the loads one thread has in flight at once). Three chunks cover
// most (pre, tile) segments of 32- to 128-neuron tiles in one round.
#ifndef BS_GATHER_SPIKES
#define BS_GATHER_SPIKES 1
#endif
#ifndef BS_GATHER_CHUNKS
#define BS_GATHER_CHUNKS 3
#endif
// 16-bit edge words: eight per chunk, so fewer chunks cover a segment
#ifndef BS_GATHER_CHUNKS16
#define BS_GATHER_CHUNKS16 2
#endif

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
// shared with the other translation units of the library
int bs_internal_fail(int code, const std::string& msg) { return fail(code, msg); }
#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(BS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

// ------------------------------------------------------------ device types
enum { FLAG_OVERFLOW = 1, FLAG_UNSTABLE = 2, FLAG_COUNT = 4, FLAG_PLASTIC = 8 };
#define BS_MAX_PL 4  // plastic projections into (and out of) one population

struct DevPop {
  int model, n, gid0, ch_begin, ch_count;
  int n_cw;          // packed count words per neuron (two 16-bit counters each)
  int n_fx;          // fixed-point channels
  int n_state;       // state variables used by the model
  int gp_begin, gp_count;  // incoming static projections (gather list)
  int tile0;         // first tile of this population
  int n_cur;         // current-input (dc_generator) channels
  // plastic projections: n_pl into this population (post side, channel
  // order), n_plpre out of it (pre side: the neurons write pre-trace history);
  // indices into the StdpDev array. n_tail: static channels summed after a
  // plastic channel of the same receptor (their values wait in LDS for the
  // update loop, which completes the sum in channel order). pl_cap: LDS
  // arrival records per sub-window.
  int n_pl, n_plpre, n_tail, pl_cap;
  int pl[BS_MAX_PL], plpre[BS_MAX_PL];
  long long tab_bytes;  // bytes of tables staged into LDS by k_window
  double P[BS_NPARAM];
};

struct DevChan {
  int kind, receptor;
  int cw, half;  // COUNT: word (0..n_cw) and half of the population's words
  int max_count; // COUNT: table holds S[0..max_count]
  int fx, shift; // FIXED: index and F
  long long table_off;
  long long lds_off;   // offset of the staged table in the LDS table area, -1 if not staged
  double weight;
  int parts, delay;    // POISSON: lam split into parts; delay in steps
  int cdf_len;
  int pre;             // POISSON: counts drawn ahead by producer blocks (pcnt)
  long long tmin, tmax;
  unsigned seed;
  int lds_len;         // COUNT: entries S[0..lds_len) staged in LDS (the rest read from the table)
  const int* gen_index;
  const double* cdf;             // POISSON: CDF table of Poisson(lam / parts)
  const unsigned short* guide;   // POISSON: 64-entry guide table
  int plastic;                   // FIXED of a plastic projection: its index in the population's pl list, else -1
  int tail;                      // static channel summed after a plastic one: its LDS tail slot, else -1
  double frac;                   // POISSON: InputDelay lerp weight of frame delay + 1 (0: whole steps)
  const long long* sched;        // SCHEDULE: [2 * cdf_len] plateau start, stop steps
  const double* sval;            // SCHEDULE: [cdf_len] plateau values
  // POISSON: counts injected by bs_net_set_drive for generator steps
  // [d0, d0 + dn) ([dn][dw], element e at e), used instead of the draws
  const int32_t* drive;
  long long d0, dn;
  int dw, n_gen;                 // n_gen: generator elements (max gen_index + 1, or the population size)
};

// One block's tile: everything its prologue needs before the population's
// blob (DevPop, channels, gather sources, staged tables) has reached LDS.
struct Tile {
  int pop, start, count, tile_in_pop;
  int ncw, nfx, ngp, nch;
  int nrec, blob_words, gp_begin, pois;  // pois: the population has a drawn-ahead channel
  int ncur, pad2;    // ncur: 1 if the population has current-input channels
  int pre_pop[4];   // pre populations of the first four gather sources
  long long gid0, blob_off, tab_bytes;  // 96 bytes: records stay 16-byte aligned
};
static_assert(sizeof(Tile) % 16 == 0, "tile records are read as 16-byte words");

// An incoming static projection of a post population. Its edges are grouped
// by (pre neuron, post tile): seg[i * (nt + 1) + t] .. seg[i * (nt + 1) + t + 1]
// are the edges from pre neuron i into tile t; an edge word holds the edge's
// offset in the post tile's LDS arrival counters, delay * RS + lane, with RS
// = (counter words or fixed-point channels of the post population) * TS the
// stride of one arrival step, so the gather adds it to a per-spike base.
struct GProj {
  int pre_pop, chan, d_min, d_max, nt;
  int rs;  // step stride of the edge words (LDS counter rows of the post population)
  long long ne;  // edges
  const uint32_t* seg;
  const uint32_t* edges;
  const long long* q;  // FIXED channels: 2^F fixed-point weights per edge
};

// Poisson channel whose counts producer blocks draw one window ahead: the
// population's neurons gid0 .. gid0 + n - 1, channel chans[ch].
struct PJob {
  long long gid0;
  int n, ch;
};

// Plastic projection (stdp_synapse.py:163-179 on event_plastic.py:454-536),
// per-edge delays d >= 1. At step n, for edge e (pre i, post j, delay d):
//   pre trace   tr_e(n) = U_i(n - d),  U_i(m) = U_i(m-1) P+ + s_i(m-1), U_i(0) = 0
//   post trace  V_j(n)  = V_j(n-1) P- + s_j(n-1)
//   post fired  s_j(n-1); the edge arrives at n iff s_i(n - 1 - d)
//   where post fired: w <- facilitate(w, tr - arrives), then
//   where it arrives: w <- depress(w, V - post fired), and the new w is
//   delivered to j at step n.
// Edges are stored grouped by (pre, post tile) like the static gather
// segments; the position in that order indexes w, la and e_pre. The CSC index
// by post lists the positions of each post neuron's incoming edges.
struct StdpDev {
  int pre_pop, post_pop, n_pre, n_post, d_min, d_max, Lu, shift, nt, pad;
  long long pre_gid0, post_gid0;
  const uint32_t* seg;      // [n_pre][nt + 1] first position of each (pre, tile) segment; [.][nt] row end
  const uint32_t* ew;       // [E] delay << 8 | lane in the post tile
  const int* e_pre;         // [E] pre neuron (local index)
  const unsigned* csc_ptr;  // [n_post + 1]
  const unsigned* csc_eid;  // [E] positions grouped by post, ascending
  double* w;                // [E]
  int* la;                  // [E] step of the edge's last arrival (-1: none yet)
  double* U;                // [Lu][n_pre] pre-trace history, slot m % Lu holds U_i(m)
  double* V;                // [n_post] post trace V_j of the last step run
  double decay_plus, decay_minus, lam, alpha, mup, mum, Wmax;
};

__device__ __forceinline__ double stdp_facilitate(const StdpDev& p, double w, double kplus) {
  const double nw = w / p.Wmax + p.lam * pow_p(1.0 - w / p.Wmax, p.mup) * kplus;
  return nw < 1.0 ? nw * p.Wmax : p.Wmax;
}
__device__ __forceinline__ double stdp_depress(const StdpDev& p, double w, double kminus) {
  const double nw = w / p.Wmax - p.alpha * p.lam * pow_p(w / p.Wmax, p.mum) * kminus;
  return nw > 0.0 ? nw * p.Wmax : 0.0;
}

struct WinArgs {
  const Tile* tiles;
  int n_tiles;               // blocks past n_tiles are Poisson producers
  unsigned short* pcnt;      // [2 * Wmax][Ntot] drawn counts, slot = step % (2 * Wmax)
  const PJob* pjobs;
  int n_pjobs;
  long long pjob_n;          // neurons over all jobs (elements per step)
  const unsigned char* blob;  // per-population blobs (BlobLayout), 16-byte aligned
  const DevPop* pops;
  const DevChan* chans;
  const void* tables;
  const GProj* gproj;
  void* state;  // [BS_NSTATE][Ntot]
  int* ref;
  long long Ntot;
  uint32_t* sb;        // spike buckets [Ls][Ntot]: local indices of the spikes of (step, pop)
  uint32_t* sb_count;  // [Ls][n_pops]
  int Ls, n_pops;
  int n0, W, Wmax;
  const long long* base_ptr;  // graph replays: n0 = *base_ptr + widx * Wmax
  int widx;
  unsigned char* spk_hist;  // [Ls][Ntot] spike flags (plastic networks), or null
  const StdpDev* stdp;      // plastic projections
  int psp;                  // plastic gather: source steps per projection and window, at most
  int max_span;        // longest gathered step range, LDS prefix size
  int TS;              // neurons per tile
  // arrival ring of events that land after their delivery window:
  // ring_cnt [Lr][NCW][Ntot] packed counts, ring_fx [Lr][NFX][Ntot] quanta
  uint32_t* ring_cnt;
  long long* ring_fx;
  int Lr, NCW, NFX;
  int H;                    // LDS counters cover arrival offsets 0 .. H-1
  int prevlen;              // steps of the previous window (its spikes are gathered now)
  const int* prevlen_ptr;   // graph replays, first window of the chunk
  unsigned long long* rec;  // (step - vm_step0) << gid_bits | gid
  int gid_bits;
  unsigned long long* rec_count;
  long long rec_cap;
  int* flags;
  const int* vm_col;
  int n_vm;
  double* vm_buf;
  long long vm_step0;
};

// ------------------------------------------------------------ window kernel
// One workgroup (256 threads) per tile of TS <= 256 neurons of one population,
// one launch per window of W steps:
//  0. stage: the population's parameters, channel descriptors, count tables
//     and Poisson CDF tables are copied into LDS;
//  1. gather: for each incoming static projection, walk the spikes emitted in
//     the steps whose deliveries can land in this window (read from the spike
//     buckets), read only the segment of each spike's row that targets this
//     tile, and add each event whose arrival step s + 1 + d falls in the
//     window into an LDS counter (integer, order free) or an LDS 2^-F
//     fixed-point accumulator. Loads are issued in phases (bucket entries,
//     then segment bounds, then edges) so each phase's loads are independent;
//  2. update: lane j < TS advances its neuron through the W steps, reading its
//     inputs from LDS, drawing the Poisson drive, applying the model step;
//     spikes are buffered in LDS;
//  3. emit: once per window, each step's spikes are appended to the
//     (step, population) bucket read by later windows and to the record list.
struct GInfo {
  const uint32_t* seg;
  const uint32_t* edges;
  const long long* q;
  int pre_gid0, nt1, kind, cw, half, fx, s_lo, rs;  // rs: the edge words' step stride
};

struct WinLayout {
  size_t cnt, fxa, spl, spn, base, pref, gtot, gslo, inp, pc, blob, end;
};

__host__ __device__ inline size_t al8(size_t x) { return (x + 7) & ~(size_t)7; }
__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }

// A population's blob, copied to LDS by every block of the population in one
// pass: its DevPop, channel records, gather-source records and staged tables.
struct BlobLayout {
  size_t pop, ch, gi, tab, end;
};
__host__ __device__ inline BlobLayout blob_layout(int nch, int ngp, size_t tab_bytes) {
  BlobLayout B;
  B.pop = 0;
  B.ch = al16(sizeof(DevPop));
  B.gi = al16(B.ch + (size_t)nch * sizeof(DevChan));
  B.tab = al16(B.gi + (size_t)ngp * sizeof(GInfo));
  B.end = al16(B.tab + tab_bytes);
  return B;
}

// LDS layout of k_window for one population (tab_bytes: staged tables).
// H: arrival offsets held in LDS counters (>= Wm).
__host__ __device__ inline WinLayout win_layout(int Wm, int H, int TS, int ncw, int nfx, int ngp, int SP,
                                                int nch, int nrec, int tsz, size_t tab_bytes, int npc) {
  WinLayout L;
  L.cnt = 0;
  L.fxa = al8(L.cnt + (size_t)H * ncw * TS * 4);
  L.spl = al8(L.fxa + (size_t)H * nfx * TS * 8);
  L.spn = al8(L.spl + (size_t)Wm * TS * 2);
  L.base = al8(L.spn + (size_t)Wm * 4);
  L.pref = al8(L.base + (size_t)Wm * 16);
  L.gtot = al8(L.pref + (size_t)ngp * SP * 4);
  L.gslo = L.gtot + (size_t)(ngp + 1) * 4;
  L.inp = al16(L.gslo + (size_t)ngp * 4);
  L.pc = al16(L.inp + (size_t)Wm * TS * nrec * tsz);
  L.blob = al16(L.pc + (size_t)Wm * TS * 2 * npc);
  L.end = L.blob + blob_layout(nch, ngp, tab_bytes).end;
  return L;
}

// Channel descriptor held in registers by the update loop.
template <typename T>
struct ChReg {
  int kind, receptor, cw, half, max_count, fx, shift, delay, parts, cdf_len, pre;
  unsigned seed;
  long long tmin, tmax;
  T weight;
  const T* tab;                 // count table S[0..tab_n) (LDS, or global when not staged)
  const T* tabg;                // the whole count table in global memory
  int tab_n;
  const double* cdf;            // Poisson CDF (LDS or global)
  const unsigned short* guide;  // Poisson guide table
  const int* gen_index;
  int plastic, tail;            // see DevChan
  double frac;                  // POISSON: lerp weight of the frame one step further back
  const long long* sched;       // SCHEDULE: plateau steps and values (global memory)
  const double* sval;
};

// pre: bit 0 the counts come from the producer blocks' ring, bit 1 counts
// injected by bs_net_set_drive cover some steps (their fields stay in the LDS
// record, read only on that path).
template <typename T>
__device__ __forceinline__ ChReg<T> load_ch(const DevChan& ch, const unsigned char* tab,
                                            const T* tables) {
  ChReg<T> c;
  c.kind = ch.kind;
  c.receptor = ch.receptor;
  c.cw = ch.cw;
  c.half = ch.half;
  c.max_count = ch.max_count;
  c.fx = ch.fx;
  c.shift = ch.shift;
  c.delay = ch.delay;
  c.parts = ch.parts;
  c.cdf_len = ch.cdf_len;
  c.pre = ch.pre | (ch.drive ? 2 : 0);
  c.seed = ch.seed;
  c.tmin = ch.tmin;
  c.tmax = ch.tmax;
  c.weight = (T)ch.weight;
  c.gen_index = ch.gen_index;
  c.plastic = ch.plastic;
  c.tail = ch.tail;
  c.frac = ch.frac;
  c.sched = ch.sched;
  c.sval = ch.sval;
  const bool staged = ch.lds_off >= 0;
  c.tab = staged ? (const T*)(tab + ch.lds_off) : tables + ch.table_off;
  c.tabg = tables + ch.table_off;
  c.tab_n = staged ? ch.lds_len : ch.max_count + 1;
  c.cdf = staged ? (const double*)(tab + ch.lds_off) : ch.cdf;
  c.guide = staged ? (const unsigned short*)((const double*)(tab + ch.lds_off) + ch.cdf_len)
                   : ch.guide;
  return c;
}

template <typename T>
struct UpdCtx {
  const WinArgs& a;
  const uint32_t* cnt;
  const long long* fxa;
  unsigned short* spl;
  int* spn;
  const DevChan* chs;
  const unsigned char* tab;
  const T* tables;
  int TS, ncw, nfx, nch, start, tid;
  bool active;
  long long gid;
  const T* inp;
  int nrec, Wm;
  const unsigned short* pc;  // [Wm][TS] drawn-ahead Poisson counts of this tile
  int ncur;                  // 1: row nrec of inp holds the summed current input
  const double* schv;        // [Wm] the current SCHEDULE channel's value per window step
};

// Value of a step_current_generator schedule at step n: the plateau with the
// last start <= n, if n is before its stop; 0 otherwise (binary search).
__device__ __forceinline__ double sched_at(const long long* se, const double* sv, int len, long long n) {
  int lo = 0, hi = len;  // first plateau with start > n
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (se[2 * mid] <= n) lo = mid + 1; else hi = mid;
  }
  return (lo > 0 && n < se[2 * (lo - 1) + 1]) ? sv[lo - 1] : 0.0;
}

// The count of generator element e at its step sgs: the injected count when
// bs_net_set_drive covers sgs, else the channel's Threefry draw.
__device__ __forceinline__ uint32_t pois_count(const int32_t* drive, long long d0, long long dn, int dw,
                                               uint32_t seed, long long sgs, int e, const double* cdf,
                                               const unsigned short* guide, int cdf_len, int parts) {
  if (drive && sgs >= d0 && sgs < d0 + dn) return (uint32_t)drive[(sgs - d0) * dw + e];
  return poisson_inv(seed, (uint32_t)sgs, (uint32_t)e, cdf, guide, cdf_len, parts);
}

// Poisson drive of neuron j at step `step`: the generator's count of step
// step - 1 - delay times the weight, or with a fractional delay the InputDelay
// lerp (delay.py:114-128) of frames d and d + 1, x_lo (1 - frac) + x_hi frac,
// then times the weight. cdf and guide point at the staged LDS copy or at
// global memory.
// DRV: the channel may carry injected counts (c.pre bit 1; dch its LDS record).
template <typename T, bool DRV>
__device__ __forceinline__ T pois_value(const ChReg<T>& c, const double* cdf, const unsigned short* guide,
                                        int j, int step, const DevChan* dch) {
  const int e = c.gen_index ? c.gen_index[j] : j;
  const long long sgs = (long long)step - 1 - c.delay;
  auto count = [&](long long sg) -> uint32_t {
    if constexpr (DRV) {
      if (c.pre & 2) return pois_count(dch->drive, dch->d0, dch->dn, dch->dw, c.seed, sg, e, cdf, guide, c.cdf_len,
                                       c.parts);
    }
    return poisson_inv(c.seed, (uint32_t)sg, (uint32_t)e, cdf, guide, c.cdf_len, c.parts);
  };
  uint32_t cn = 0;
  if (e >= 0 && sgs >= 0 && sgs > c.tmin && sgs <= c.tmax) cn = count(sgs);
  if (c.frac == 0.0) return (T)cn * c.weight;
  const long long sgh = sgs - 1;
  uint32_t ch = 0;
  if (e >= 0 && sgh >= 0 && sgh > c.tmin && sgh <= c.tmax) ch = count(sgh);
  const T f = (T)c.frac;
  return ((T)cn * ((T)1 - f) + (T)ch * f) * c.weight;
}

// Input of one channel at window step w for lane tid (neuron j).
template <typename T>
__device__ __forceinline__ T chan_value(const UpdCtx<T>& u, const ChReg<T>& c, int ci, int w, int step,
                                        int lane) {
  if (c.kind == BS_CH_COUNT) {
    uint32_t k = (u.cnt[(w * u.ncw + c.cw) * u.TS + lane] >> (16 * c.half)) & 0xFFFFu;
    if (k > (uint32_t)c.max_count) {  // cannot happen for consistent edges
      atomicOr(u.a.flags, FLAG_COUNT);
      k = (uint32_t)c.max_count;
    }
    return k < (uint32_t)c.tab_n ? c.tab[k] : c.tabg[k];
  }
  if (c.kind == BS_CH_POISSON) {
    if (c.pre & 1) return (T)u.pc[w * u.TS + lane] * c.weight;
    return pois_value<T, true>(c, c.cdf, c.guide, u.start + lane, step, &u.chs[ci]);
  }
  if (c.kind == BS_CH_CURRENT) {  // dc_generator: amplitude * weight inside [tmin, tmax)
    const int j = u.start + lane;
    const int e = c.gen_index ? c.gen_index[j] : j;
    return (e >= 0 && (long long)step >= c.tmin && (long long)step < c.tmax) ? c.weight : (T)0;
  }
  if (c.kind == BS_CH_SCHEDULE) {  // step_current_generator: staged per window step
    const int j = u.start + lane;
    const int e = c.gen_index ? c.gen_index[j] : j;
    return e >= 0 ? (T)u.schv[w] : (T)0;
  }
  if (c.kind == BS_CH_AC) {  // ac_generator: (A sin(omega t + phi) + offset) * weight
    const int j = u.start + lane;
    const int e = c.gen_index ? c.gen_index[j] : j;
    if (e < 0 || (long long)step < c.tmin || (long long)step >= c.tmax) return (T)0;
    const double* p = c.cdf;  // {A, omega, phi, offset, weight, dt}
    const double t = (double)step * p[5];
    return (T)((p[0] * sin_p(p[1] * t + p[2]) + p[3]) * p[4]);
  }
  const long long q = u.fxa[(w * u.nfx + c.fx) * u.TS + lane];  // BS_CH_FIXED
  return (T)__builtin_ldexp((double)q, -c.shift);
}

// ---- a lane's neuron over the window -----------------------------------
// run(wa, wb, in, out) advances the lane's neuron through window steps
// wa .. wb - 1: in(w, a0, a1, cur) gives the summed inputs of step w (read
// after the step's integration, before they enter the state), out(w, spike,
// v) takes the step's outcome (v = V_m after it); both are called once per
// step, in step order.

// Models whose step is one straight computation, f(s, r, a0, a1, cur).
template <typename T, typename F>
struct StepLane {
  T s[BS_NSTATE];
  int r;
  int ns;
  const F& f;
  __device__ __forceinline__ StepLane(int ns_, const F& f_) : r(0), ns(ns_), f(f_) {
#pragma unroll
    for (int k = 0; k < BS_NSTATE; ++k) s[k] = (T)0;
  }
  __device__ __forceinline__ void load(const WinArgs& a, long long gid) {
    const T* st = (const T*)a.state;
#pragma unroll
    for (int k = 0; k < BS_NSTATE; ++k)
      if (k < ns) s[k] = st[(long long)k * a.Ntot + gid];
    r = a.ref[gid];
  }
  __device__ __forceinline__ void store(const WinArgs& a, long long gid) const {
    T* st = (T*)a.state;
#pragma unroll
    for (int k = 0; k < BS_NSTATE; ++k)
      if (k < ns) st[(long long)k * a.Ntot + gid] = s[k];
    a.ref[gid] = r;
  }
  template <typename IN, typename OUT>
  __device__ __forceinline__ void run(int wa, int wb, IN&& in, OUT&& out) {
    for (int w = wa; w < wb; ++w) {
      T a0, a1, cur;
      in(w, a0, a1, cur);
      const bool spike = f(s, r, a0, a1, cur);
      out(w, spike, (double)s[0]);
    }
  }
};

// aeif_cond_alpha: each lane runs its own substeps and moves to its next step
// as soon as its current one is integrated, so a wave takes as many substep
// iterations as its slowest lane needs over the whole window, not, at every
// step, as many as that step's slowest lane. Per lane the operations and
// their order are those of step_aeif.
struct AeifRun {
  AeifLane L;
  int r;
  int unstable;
  const double* P;
  AeifR R;
  __device__ __forceinline__ explicit AeifRun(const double* P_) : r(0), unstable(0), P(P_), R(aeif_recips(P_)) {}
  __device__ __forceinline__ void load(const WinArgs& a, long long gid) {
    const double* st = (const double*)a.state;
    double s[BS_NSTATE];
#pragma unroll
    for (int k = 0; k < BS_NSTATE; ++k) s[k] = st[(long long)k * a.Ntot + gid];
    aeif_load(L, s);
    r = a.ref[gid];
  }
  __device__ __forceinline__ void store(const WinArgs& a, long long gid) const {
    double* st = (double*)a.state;
    double s[BS_NSTATE];
    aeif_store(L, s);
#pragma unroll
    for (int k = 0; k < BS_NSTATE; ++k) st[(long long)k * a.Ntot + gid] = s[k];
    a.ref[gid] = r;
  }
  template <typename IN, typename OUT>
  __device__ __forceinline__ void run(int wa, int wb, IN&& in, OUT&& out) {
    int w = wa;
    bool live = wa < wb;
    if (live) {
      aeif_begin(P, L);
      aeif_rhs(P, R, L.y, r > 0, L.istim, L.k0);
    }
    // (finishing a step as soon as it is integrated measured faster than
    // holding finished lanes until a quarter or half of the wave is done)
    while (live) {
      bool need = false;  // k0 is stale
      if (aeif_more(P, L)) need = aeif_substep(P, R, L, r, &unstable) && L.tl < P[C_DT];
      if (!aeif_more(P, L)) {  // step w integrated
        double a0, a1, cur;
        in(w, a0, a1, cur);
        aeif_finish(P, L, r, a0, a1, cur);
        out(w, L.spiked, L.y.V);
        need = ++w < wb;
        if (need) aeif_begin(P, L);
        else live = false;
      }
      if (need) aeif_rhs(P, R, L.y, r > 0, L.istim, L.k0);
    }
  }
};

// The window of one lane: inputs summed in channel order, model step,
// voltmeter tap, spike buffered in LDS. HAS_CUR: the population has current
// inputs (row nrec of inp); without them the model sees a zero current.
template <typename T, bool HAS_CUR, typename Lane>
__device__ __forceinline__ void update_loop(const UpdCtx<T>& u, Lane& m) {
  const WinArgs& a = u.a;
  if (!u.active) return;
  m.load(a, u.gid);
  const int vmc = a.n_vm > 0 ? a.vm_col[u.gid] : -1;
  unsigned smask = 0u;  // bit w: the neuron fired at step n0 + w
  m.run(
      0, a.W,
      [&](int w, T& acc0, T& acc1, T& cur) __attribute__((always_inline)) {
        acc0 = u.inp[w * u.TS + u.tid];
        acc1 = u.nrec > 1 ? u.inp[(u.Wm + w) * u.TS + u.tid] : (T)0;
        cur = (T)0;
        if constexpr (HAS_CUR) cur = u.inp[(u.nrec * u.Wm + w) * u.TS + u.tid];
      },
      [&](int w, bool spike, double v) __attribute__((always_inline)) {
        const int step = a.n0 + w;
        if (vmc >= 0) a.vm_buf[(long long)(step - a.vm_step0) * a.n_vm + vmc] = v;
        if (spike) {
          smask |= 1u << w;
          const int p = atomicAdd(&u.spn[w], 1);
          u.spl[w * u.TS + p] = (unsigned short)u.tid;
        }
      });
  if (a.spk_hist) {  // (read from the next window on)
    const int hb = a.n0 % a.Ls;
    for (int w = 0; w < a.W; ++w) {
      const int hs = hb + w;
      a.spk_hist[(long long)(hs < a.Ls ? hs : hs - a.Ls) * a.Ntot + u.gid] = (smask >> w) & 1u;
    }
  }
  m.store(a, u.gid);
}

// ------------------------------------------------------------ plastic path
// LDS of a population with plastic inputs, after the window layout: bucket
// prefix per plastic projection over its source steps, per-(step, projection,
// lane) arrival offsets (scanned in place, then cursors), scan partials, the
// facilitation masks and the list of lanes that fired, the tail channel
// values, and the arrival records (position, pre trace) of one sub-window.
struct PlLayout {
  size_t ppref, poff, part, fm, fl, tail, rk, rtr, end;
};
__host__ __device__ inline PlLayout pl_layout(size_t base, int Wm, int TS, int npl, int psp, int ntail, int tsz,
                                              int cap) {
  PlLayout L;
  L.ppref = al16(base);
  L.poff = al16(L.ppref + (size_t)npl * (psp + 1) * 4);
  L.part = al16(L.poff + ((size_t)Wm * npl * TS + 1) * 4);
  L.fm = al16(L.part + 256 * 4);
  L.fl = al16(L.fm + (size_t)TS * 4);
  L.tail = al16(L.fl + ((size_t)TS + 1) * 4);
  L.rk = al16(L.tail + (size_t)ntail * Wm * TS * tsz);
  L.rtr = al16(L.rk + (size_t)cap * 4);
  L.end = al16(L.rtr + (size_t)cap * 8);
  return L;
}

struct PlCtx {
  int* ppref;
  int* poff;
  int* part;
  unsigned* fm;
  int* fl;
  const void* tail;  // T [n_tail][Wm][TS]
  uint32_t* rk;
  double* rtr;
  int npl, nplpre, cap, S;  // S = npl * TS keys per window step
  int tile_in_pop;
};

// Bucket sizes of the source steps of every plastic projection into this
// population: a spike of step s arrives at s + 1 + d, inside the window for
// s in [n0 - 1 - d_max, n0 + W - 2 - d_min]; inclusive prefix per projection.
__device__ __forceinline__ void pl_prefix(const WinArgs& a, const DevPop& pop, const PlCtx& c, int tid) {
  const int P1 = a.psp + 1;
  for (int k = tid; k < c.npl * P1; k += 256) {
    const int p = k / P1, t = k - p * P1;
    int v = 0;
    if (t > 0) {
      const StdpDev& P = a.stdp[pop.pl[p]];
      const int s_lo = max(0, a.n0 - 1 - P.d_max), s_hi = a.n0 + a.W - 2 - P.d_min;
      const int s = s_lo + t - 1;
      if (s <= s_hi) v = (int)a.sb_count[(s % a.Ls) * a.n_pops + P.pre_pop];
    }
    c.ppref[k] = v;
  }
  __syncthreads();
  if (tid < c.npl) {
    int acc = 0;
    for (int t = 1; t < P1; ++t) {
      acc += c.ppref[tid * P1 + t];
      c.ppref[tid * P1 + t] = acc;
    }
  }
  __syncthreads();
}

// One walk over the plastic arrivals of window steps [wa, wb): MODE 0 counts
// them per key (step, projection, lane); MODE 1 places each record at its
// key's cursor. One spike per thread; a (pre, tile) segment holds a few edges.
template <int MODE>
__device__ __forceinline__ void pl_pass(const WinArgs& a, const DevPop& pop, const PlCtx& c, int TS, int wa, int wb, int tid) {
  const int P1 = a.psp + 1;
  for (int p = 0; p < c.npl; ++p) {
    const StdpDev& P = a.stdp[pop.pl[p]];
    const int* pr = c.ppref + p * P1;
    const int tot = pr[P1 - 1];
    const int s_lo = max(0, a.n0 - 1 - P.d_max);
    for (int idx = tid; idx < tot; idx += 256) {
      int lo = 0, hi = P1 - 1;  // pr[lo] <= idx < pr[lo + 1]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pr[mid] <= idx) lo = mid; else hi = mid;
      }
      const int s = s_lo + lo;
      const int i = (int)a.sb[(long long)(s % a.Ls) * a.Ntot + P.pre_gid0 + (idx - pr[lo])];
      const uint32_t* sg = P.seg + (long long)i * (P.nt + 1) + c.tile_in_pop;
      const uint32_t k0 = sg[0], k1 = sg[1];
      double tr = 0.0;
      if (MODE == 1 && k1 > k0) tr = P.U[(long long)((s + 1) % P.Lu) * P.n_pre + i];
      for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t word = P.ew[k];
        const int w = s + 1 + (int)(word >> 8) - a.n0;
        if (w < wa || w >= wb) continue;
        const int key = (w * c.npl + p) * TS + (int)(word & 255u);
        if (MODE == 0) {
          atomicAdd(&c.poff[key], 1);
        } else {
          const int pos = atomicAdd(&c.poff[key], 1);
          if (pos < c.cap) {
            c.rk[pos] = k;
            c.rtr[pos] = tr;
          }
        }
      }
    }
  }
}

// In-place exclusive scan of x[0 .. M), x[M] = total (256 threads).
__device__ __forceinline__ void block_exscan(int* x, int M, int* part, int tid) {
  const int per = (M + 255) / 256;
  const int b = min(tid * per, M), e = min(b + per, M);
  int sum = 0;
  for (int k = b; k < e; ++k) sum += x[k];
  part[tid] = sum;
  __syncthreads();
  if (tid < 64) {  // one wave: four partials per lane, then a shuffle scan
    int v[4], t = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = part[tid * 4 + q];
      t += v[q];
    }
    int incl = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (tid >= o) incl += y;
    }
    int run = incl - t;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      part[tid * 4 + q] = run;
      run += v[q];
    }
    if (tid == 63) x[M] = incl;
  }
  __syncthreads();
  int run = part[tid];
  for (int k = b; k < e; ++k) {
    const int t = x[k];
    x[k] = run;
    run += t;
  }
  __syncthreads();
}

// Facilitations an edge still owes: post spikes at f - 1 for the window steps
// f = n0 + b, b in mask (ascending), each with the pre trace U_i(f - d) and no
// arrival at f (k+ = trace - 0).
__device__ __forceinline__ double stdp_pending(const StdpDev& P, double w, unsigned long long m, int n0, int i,
                                               int d) {
  while (m) {
    const int b = __builtin_ctzll(m);
    m &= m - 1ull;
    const int mu = n0 + b - d;
    const double tr = mu >= 0 ? P.U[(long long)(mu % P.Lu) * P.n_pre + i] : 0.0;
    w = stdp_facilitate(P, w, tr - 0.0);
  }
  return w;
}

// The W-step loop of a population with plastic inputs or outputs, run by all
// 256 threads (block barriers between sub-windows). Lane j < TS advances its
// neuron as update_loop does; in addition, at each step it updates its post
// traces, processes the plastic edges arriving at it (pending facilitations,
// facilitate/depress, deposit) and completes the channel sum in channel
// order, and it writes its pre-trace history. After the loop the block
// facilitates the incoming edges of the neurons that fired and that no later
// arrival of the window has brought up to date.
template <typename T, bool HAS_CUR, typename Lane>
__device__ __forceinline__ void update_loop_pl(const UpdCtx<T>& u, const PlCtx& c, const DevPop& pop, Lane& m) {
  const WinArgs& a = u.a;
  const int tid = u.tid, TS = u.TS;
  const bool act = u.active && tid < TS;
  double vtr[BS_MAX_PL], utr[BS_MAX_PL];
#pragma unroll
  for (int q = 0; q < BS_MAX_PL; ++q) vtr[q] = utr[q] = 0.0;
  unsigned fmask = 0u;  // bit b: the neuron fired at step n0 + b - 1
  unsigned smask = 0u;  // bit w: the neuron fired at step n0 + w
  const int il = u.start + tid;  // local index in the population
  int vmc = -1;
  if (act) {
    m.load(a, u.gid);
    vmc = a.n_vm > 0 ? a.vm_col[u.gid] : -1;
    if (a.n0 >= 1 && a.spk_hist[(long long)((a.n0 - 1) % a.Ls) * a.Ntot + u.gid]) fmask = 1u;
    // (constant indices into the trace arrays keep them in registers)
#pragma unroll
    for (int p = 0; p < BS_MAX_PL; ++p)
      if (p < c.npl) vtr[p] = a.stdp[pop.pl[p]].V[il];
#pragma unroll
    for (int q = 0; q < BS_MAX_PL; ++q)
      if (q < c.nplpre) {
        const StdpDev& P = a.stdp[pop.plpre[q]];
        utr[q] = P.U[(long long)(a.n0 % P.Lu) * P.n_pre + il];
      }
  }
  if (c.npl > 0) {  // arrival counts of the window, scanned into offsets
    const int M = a.W * c.S;
    for (int k = tid; k <= M; k += 256) c.poff[k] = 0;
    pl_prefix(a, pop, c, tid);  // (its barriers also order the zeroing)
    pl_pass<0>(a, pop, c, TS, 0, a.W, tid);
    __syncthreads();
    block_exscan(c.poff, M, c.part, tid);
  }
  const T* tail = (const T*)c.tail;
  // ring slots of window step 0 (step w: base + w, wrapped once; both rings
  // are longer than a window), so that no step pays an integer division
  const int hb = a.n0 % a.Ls;
  int ub[BS_MAX_PL];
#pragma unroll
  for (int q = 0; q < BS_MAX_PL; ++q) ub[q] = q < c.nplpre ? (a.n0 + 1) % a.stdp[pop.plpre[q]].Lu : 0;
  // the channels a step completes (plastic ones and the tail channels after
  // them), as a mask over the first 64; any beyond are scanned
  unsigned long long chm = 0ull;
  if (c.npl > 0)
    for (int ch = 0; ch < u.nch && ch < 64; ++ch)
      if (u.chs[ch].plastic >= 0 || u.chs[ch].tail >= 0) chm |= 1ull << ch;
  // per-step constants of the plastic projections, loaded once per window
  double dmn[BS_MAX_PL], dpl[BS_MAX_PL];
  double* Uq[BS_MAX_PL];  // this neuron's column of each pre-trace history
  int Luq[BS_MAX_PL], npq[BS_MAX_PL];
#pragma unroll
  for (int q = 0; q < BS_MAX_PL; ++q) {
    dmn[q] = q < c.npl ? a.stdp[pop.pl[q]].decay_minus : 0.0;
    dpl[q] = 0.0;
    Uq[q] = nullptr;
    Luq[q] = 1;
    npq[q] = 0;
    if (q < c.nplpre) {
      const StdpDev& P = a.stdp[pop.plpre[q]];
      dpl[q] = P.decay_plus;
      Uq[q] = P.U + il;
      Luq[q] = P.Lu;
      npq[q] = P.n_pre;
    }
  }
  for (int wa = 0; wa < a.W;) {
    int wb = a.W;
    if (c.npl > 0) {  // the longest run of steps whose records fit (at least one step)
      const int base = c.poff[wa * c.S];
      if (c.poff[(wa + 1) * c.S] - base > c.cap && tid == 0) atomicOr(a.flags, FLAG_PLASTIC);
      wb = wa + 1;
      while (wb < a.W && c.poff[(wb + 1) * c.S] - base <= c.cap) ++wb;
      __syncthreads();
      for (int k = wa * c.S + tid; k < wb * c.S; k += 256) c.poff[k] -= base;
      __syncthreads();
      pl_pass<1>(a, pop, c, TS, wa, wb, tid);
      __syncthreads();
    }
    if (act) {
      m.run(
          wa, wb,
          [&](int w, T& acc0, T& acc1, T& cur) __attribute__((always_inline)) {
            const int step = a.n0 + w;
            acc0 = u.inp[w * TS + tid];
            acc1 = u.nrec > 1 ? u.inp[(u.Wm + w) * TS + tid] : (T)0;
            cur = (T)0;
            if constexpr (HAS_CUR) cur = u.inp[(u.nrec * u.Wm + w) * TS + tid];
            const bool post_fired = (fmask >> w) & 1u;
            if (c.npl > 0) {
#pragma unroll
              for (int p = 0; p < BS_MAX_PL; ++p)
                if (p < c.npl) vtr[p] = vtr[p] * dmn[p] + (post_fired ? 1.0 : 0.0);
              // the channels from the first plastic one of each receptor, in order
              const auto chan = [&](int ch) __attribute__((always_inline)) {
                const DevChan& dc = u.chs[ch];
                T v;
                if (dc.plastic >= 0) {
                  const int p = dc.plastic;
                  const StdpDev& P = a.stdp[pop.pl[p]];
                  // every trace loaded, then a select on the values: a load
                  // under each branch would be merged into one load at index
                  // p, which moves the trace array to scratch memory
                  static_assert(BS_MAX_PL == 4, "trace select below");
                  const double v0 = vtr[0], v1 = vtr[1], v2 = vtr[2], v3 = vtr[3];
                  const double vt = p == 0 ? v0 : (p == 1 ? v1 : (p == 2 ? v2 : v3));
                  const int key = (w * c.npl + p) * TS + tid;
                  const int b = key == wa * c.S ? 0 : min(c.poff[key - 1], c.cap);
                  const int e = min(c.poff[key], c.cap);
                  long long dep = 0;
                  for (int q = b; q < e; ++q) {
                    const uint32_t k = c.rk[q];
                    double wv = P.w[k];
                    const int la = P.la[k];
                    const int i = P.e_pre[k];
                    const int d = (int)(P.ew[k] >> 8);
                    // facilitations since the edge's last arrival (or the window start)
                    const int lo = la > a.n0 - 1 ? la : a.n0 - 1;
                    const unsigned long long mk =
                        (unsigned long long)fmask & ((1ull << w) - 1ull) & ~((1ull << (lo - a.n0 + 1)) - 1ull);
                    wv = stdp_pending(P, wv, mk, a.n0, i, d);
                    if (post_fired) wv = stdp_facilitate(P, wv, c.rtr[q] - 1.0);
                    wv = stdp_depress(P, wv, vt - (post_fired ? 1.0 : 0.0));
                    P.w[k] = wv;
                    P.la[k] = step;
                    dep += (long long)__builtin_rint(__builtin_ldexp(wv, P.shift));
                  }
                  v = (T)__builtin_ldexp((double)dep, -P.shift);
                } else if (dc.tail >= 0) {
                  v = tail[(dc.tail * u.Wm + w) * TS + tid];
                } else {
                  return;
                }
                if (dc.receptor == 0) acc0 = acc0 + v;
                else acc1 = acc1 + v;
              };
              for (unsigned long long mm = chm; mm; mm &= mm - 1ull) chan(__builtin_ctzll(mm));
              for (int ch = 64; ch < u.nch; ++ch) chan(ch);
            }
          },
          [&](int w, bool spike, double vm) __attribute__((always_inline)) {
            const int step = a.n0 + w;
            if (vmc >= 0) a.vm_buf[(long long)(step - a.vm_step0) * a.n_vm + vmc] = vm;
            if (spike) {
              smask |= 1u << w;
              if (w + 1 < a.W) fmask |= 1u << (w + 1);
              const int pp = atomicAdd(&u.spn[w], 1);
              u.spl[w * TS + pp] = (unsigned short)tid;
            }
          });
    }
    if (c.npl > 0) __syncthreads();
    wa = wb;
  }
  if (act) {
    // the window's spike history and pre-trace history, written after the
    // steps (nothing in this window reads them: arrivals read pre traces of
    // steps up to n0, and the spike history is read from the next window on)
    for (int w = 0; w < a.W; ++w) {
      const int hs = hb + w;
      a.spk_hist[(long long)(hs < a.Ls ? hs : hs - a.Ls) * a.Ntot + u.gid] = (smask >> w) & 1u;
    }
#pragma unroll
    for (int q = 0; q < BS_MAX_PL; ++q)
      if (q < c.nplpre) {
        double t = utr[q];
        for (int w = 0; w < a.W; ++w) {
          t = t * dpl[q] + (((smask >> w) & 1u) ? 1.0 : 0.0);
          const int us = ub[q] + w;  // slot of step n0 + w + 1
          Uq[q][(long long)(us < Luq[q] ? us : us - Luq[q]) * npq[q]] = t;
        }
      }
    m.store(a, u.gid);
#pragma unroll
    for (int p = 0; p < BS_MAX_PL; ++p)
      if (p < c.npl) a.stdp[pop.pl[p]].V[il] = vtr[p];
  }
  if (c.npl == 0) return;
  // facilitation of the incoming edges of the neurons that fired
  if (tid < TS) c.fm[tid] = act ? fmask : 0u;
  if (tid == 0) c.fl[TS] = 0;
  __syncthreads();
  if (tid < TS && c.fm[tid]) c.fl[atomicAdd(&c.fl[TS], 1)] = tid;
  __syncthreads();
  const int nl = c.fl[TS];
  for (int p = 0; p < c.npl; ++p) {
    const StdpDev& P = a.stdp[pop.pl[p]];
    for (int li = 0; li < nl; ++li) {
      const int lane = c.fl[li];
      const unsigned fmk = c.fm[lane];
      const int j = u.start + lane;
      const unsigned c0 = P.csc_ptr[j], c1 = P.csc_ptr[j + 1];
      for (unsigned cc = c0 + tid; cc < c1; cc += 256) {
        const uint32_t k = P.csc_eid[cc];
        const int la = P.la[k];
        const int lo = la > a.n0 - 1 ? la : a.n0 - 1;
        const unsigned long long m = (unsigned long long)fmk & ~((1ull << (lo - a.n0 + 1)) - 1ull);
        if (!m) continue;
        P.w[k] = stdp_pending(P, P.w[k], m, a.n0, P.e_pre[k], (int)(P.ew[k] >> 8));
      }
    }
  }
}

template <typename T, bool HAS_CUR, typename Lane>
__device__ __forceinline__ void run_update(const UpdCtx<T>& u, bool plastic, const PlCtx& pl, const DevPop& pop,
                                           Lane& m) {
  if (plastic)
    update_loop_pl<T, HAS_CUR>(u, pl, pop, m);
  else if (u.tid < u.TS)
    update_loop<T, HAS_CUR>(u, m);
}

// Producer blocks: draw the Poisson counts of steps n1 .. n1 + Wmax - 1 for
// every drawn-ahead channel (the next window's inputs) while the tile blocks
// run the current window on the other CUs. The counts are functions of
// (seed, step, element) only, so they equal what the inputs phase would draw.
__device__ __forceinline__ void produce_poisson(const WinArgs& a, int n1, int blk, int nblk) {
  const long long per = a.pjob_n;
  const long long total = per * a.Wmax;
  const int L2 = 2 * a.Wmax;
  for (long long idx = (long long)blk * 256 + threadIdx.x; idx < total; idx += (long long)nblk * 256) {
    const int w = (int)(idx / per);
    long long r = idx - (long long)w * per;
    int j = 0;
    while (j + 1 < a.n_pjobs && r >= a.pjobs[j].n) {
      r -= a.pjobs[j].n;
      ++j;
    }
    const PJob J = a.pjobs[j];
    const DevChan& c = a.chans[J.ch];
    const int step = n1 + w;
    const int i = (int)r;
    const int e = c.gen_index ? c.gen_index[i] : i;
    const long long sgs = (long long)step - 1 - c.delay;
    uint32_t cn = 0;
    if (e >= 0 && sgs >= 0 && sgs > c.tmin && sgs <= c.tmax)
      cn = pois_count(c.drive, c.d0, c.dn, c.dw, c.seed, sgs, e, c.cdf, c.guide, c.cdf_len, c.parts);
    a.pcnt[(long long)(step % L2) * a.Ntot + J.gid0 + i] = (unsigned short)cn;
  }
}

// The counts of a run's first window (later windows get theirs from the
// producer blocks of the launch before).
__global__ __launch_bounds__(256) void k_pois_fill(WinArgs a) {
  produce_poisson(a, a.n0, (int)blockIdx.x, (int)gridDim.x);
}

// PLASTIC: the network has plastic projections (a separate instantiation, so the
// static kernel carries none of the plastic path's registers). NARROW: the
// static edge words are 16-bit (every projection's delay * stride + lane fits).
template <typename T, bool PLASTIC, bool NARROW>
__device__ __forceinline__ void window_body(WinArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  if (a.base_ptr) {
    const long long b = *a.base_ptr;
    a.n0 = (int)b + a.widx * a.Wmax;
    a.vm_step0 = b;
  }
  if ((int)blockIdx.x >= a.n_tiles) {  // spare CUs: the next window's Poisson counts
    produce_poisson(a, a.n0 + a.W, (int)blockIdx.x - a.n_tiles, (int)gridDim.x - a.n_tiles);
    return;
  }
  // XCD-aware tile order: the dispatcher deals blocks round-robin over the
  // eight XCDs (block b runs on XCD b % 8), so each XCD is given one
  // contiguous run of tiles. A spike's edge row is laid out tile-major, and
  // every XCD then fetches only its own slice of the row into its L2.
  const int tix = [&a] {
    const int G = a.n_tiles, x = (int)(blockIdx.x & 7u), k = (int)(blockIdx.x >> 3);
    const int q = G >> 3, rem = G & 7;
    return x * q + min(x, rem) + k;
  }();
  const Tile tl = a.tiles[tix];
  const int tid = (int)threadIdx.x;
  const int TS = a.TS;
  const int tile_in_pop = tl.tile_in_pop;
  const int W = a.W, Wm = a.Wmax;
  const int ncw = tl.ncw, nfx = tl.nfx, ngp = tl.ngp, nch = tl.nch;
  const int SP = a.max_span + 1;
  const int nrec = tl.nrec;
  const int H = a.H;
  const WinLayout Lo = win_layout(Wm, H, TS, ncw, nfx, ngp, SP, nch, nrec + tl.ncur, (int)sizeof(T),
                                  tl.tab_bytes, tl.pois);
  const BlobLayout Bl = blob_layout(nch, ngp, tl.tab_bytes);
  uint32_t* cnt = (uint32_t*)(lds + Lo.cnt);
  long long* fxa = (long long*)(lds + Lo.fxa);
  unsigned short* spl = (unsigned short*)(lds + Lo.spl);
  int* spn = (int*)(lds + Lo.spn);
  unsigned long long* base = (unsigned long long*)(lds + Lo.base);
  int* pref = (int*)(lds + Lo.pref);
  int* gtot = (int*)(lds + Lo.gtot);
  int* gslo = (int*)(lds + Lo.gslo);
  unsigned char* blob = lds + Lo.blob;
  const DevPop& gpop = *(const DevPop*)(blob + Bl.pop);
  const double* P = gpop.P;
  DevChan* chs = (DevChan*)(blob + Bl.ch);
  const GInfo* gi = (const GInfo*)(blob + Bl.gi);
  unsigned char* tab = blob + Bl.tab;
  T* inp = (T*)(lds + Lo.inp);  // [nrec][Wm][TS] summed inputs per step and lane
  unsigned short* pc = (unsigned short*)(lds + Lo.pc);

  // ---- 0. stage: one level of loads after the tile record -- the blob, the
  // drawn-ahead Poisson counts, the ring slots of this window, the bucket
  // counts the gather needs. Every load is issued into registers before the
  // first LDS store waits on one, so the stage costs one memory round trip
  // (the register slots cover the common sizes; loops take the rest).
  const long long tile_gid0 = tl.gid0 + tl.start;
  const int PL = (a.base_ptr && a.widx == 0) ? *a.prevlen_ptr : a.prevlen;
  const int lgTS = TS == 256 ? 8 : (TS == 128 ? 7 : (TS == 64 ? 6 : 5));
  const int lane_t = tid & (TS - 1), row0 = tid >> lgTS, rows_per = 256 >> lgTS;
  const bool lane_ok = lane_t < tl.count;
  const int L2 = 2 * Wm;
  constexpr int kRB = 2, kRP = 4, kRC = 8, kRF = 4;
  const uint4* bsrc = (const uint4*)(a.blob + tl.blob_off);
  uint4 rb[kRB];
  unsigned short rp[kRP];
  uint32_t rc[kRC];
  long long rf[kRF];
  int rq = 0;
#pragma unroll
  for (int i = 0; i < kRB; ++i) {
    const int k = tid + 256 * i;
    rb[i] = k < tl.blob_words ? bsrc[k] : make_uint4(0u, 0u, 0u, 0u);
  }
#pragma unroll
  for (int i = 0; i < kRP; ++i) {  // pc[w][lane]: row = step w
    const int w = row0 + rows_per * i;
    rp[i] = (tl.pois && w < W && lane_ok)
                ? a.pcnt[(long long)((a.n0 + w) % L2) * a.Ntot + tile_gid0 + lane_t]
                : (unsigned short)0;
  }
#pragma unroll
  for (int i = 0; i < kRC; ++i) {  // cnt[w][cw][lane], w < W (no rows when ncw = 0)
    const int r = row0 + rows_per * i, w = ncw > 0 ? r / ncw : W, cw = r - w * ncw;
    rc[i] = (w < W && lane_ok)
                ? a.ring_cnt[((size_t)((a.n0 + w) % a.Lr) * a.NCW + cw) * a.Ntot + tile_gid0 + lane_t]
                : 0u;
  }
#pragma unroll
  for (int i = 0; i < kRF; ++i) {  // fxa[w][fx][lane], w < W
    const int r = row0 + rows_per * i;
    const int w = nfx > 0 ? r / nfx : W, fx = r - w * nfx;
    rf[i] = (nfx > 0 && w < W && lane_ok)
                ? a.ring_fx[((size_t)((a.n0 + w) % a.Lr) * a.NFX + fx) * a.Ntot + tile_gid0 + lane_t]
                : 0ll;
  }
  // bucket sizes of every (projection, step) the gather reads: the spikes of
  // the previous window, each delivered once with all of its delays
  if (tid < ngp * SP) {
    const int g = tid / SP, t = tid - g * SP;
    // constant indices only: a runtime index into tl would move it to scratch
    const int pre_pop = g == 0   ? tl.pre_pop[0]
                        : g == 1 ? tl.pre_pop[1]
                        : g == 2 ? tl.pre_pop[2]
                        : g == 3 ? tl.pre_pop[3]
                                 : a.gproj[tl.gp_begin + g].pre_pop;
    const int s = a.n0 - PL + t;
    rq = (t > 0 && t <= PL) ? (int)a.sb_count[((s - 1) % a.Ls) * a.n_pops + pre_pop] : 0;
  }
  if (tix == 0) {  // clear the bucket counts of the next window's steps
    for (int t = tid; t < Wm * a.n_pops; t += 256) {
      const int st = a.n0 + W + t / a.n_pops;
      a.sb_count[(st % a.Ls) * a.n_pops + t % a.n_pops] = 0u;
    }
  }
  // now the stores: LDS copies, and the ring slots taken out are zeroed
  {
    uint4* dst = (uint4*)blob;
#pragma unroll
    for (int i = 0; i < kRB; ++i) {
      const int k = tid + 256 * i;
      if (k < tl.blob_words) dst[k] = rb[i];
    }
    for (int k = tid + 256 * kRB; k < tl.blob_words; k += 256) dst[k] = bsrc[k];
  }
  if (tl.pois) {
#pragma unroll
    for (int i = 0; i < kRP; ++i) {
      const int w = row0 + rows_per * i;
      if (w < W) pc[w * TS + lane_t] = rp[i];
    }
    for (int w = row0 + rows_per * kRP; w < W; w += rows_per)
      pc[w * TS + lane_t] = lane_ok ? a.pcnt[(long long)((a.n0 + w) % L2) * a.Ntot + tile_gid0 + lane_t]
                                    : (unsigned short)0;
  }
  // events already deposited for this window (rows w < W), the rest zero
#pragma unroll
  for (int i = 0; i < kRC; ++i) {
    const int r = row0 + rows_per * i, w = ncw > 0 ? r / ncw : W, cw = r - w * ncw;
    if (r < H * ncw) {
      if (w < W && lane_ok)
        a.ring_cnt[((size_t)((a.n0 + w) % a.Lr) * a.NCW + cw) * a.Ntot + tile_gid0 + lane_t] = 0u;
      cnt[r * TS + lane_t] = rc[i];
    }
  }
  for (int r = row0 + rows_per * kRC; r < H * ncw; r += rows_per) {
    const int w = r / ncw, cw = r - w * ncw;
    uint32_t v = 0u;
    if (w < W && lane_ok) {
      const size_t o = ((size_t)((a.n0 + w) % a.Lr) * a.NCW + cw) * a.Ntot + tile_gid0 + lane_t;
      v = a.ring_cnt[o];
      a.ring_cnt[o] = 0u;
    }
    cnt[r * TS + lane_t] = v;
  }
#pragma unroll
  for (int i = 0; i < kRF; ++i) {
    const int r = row0 + rows_per * i;
    if (r < H * nfx) {
      const int w = r / nfx, fx = r - w * nfx;
      if (w < W && lane_ok)
        a.ring_fx[((size_t)((a.n0 + w) % a.Lr) * a.NFX + fx) * a.Ntot + tile_gid0 + lane_t] = 0;
      fxa[r * TS + lane_t] = rf[i];
    }
  }
  for (int r = row0 + rows_per * kRF; r < H * nfx; r += rows_per) {
    const int w = r / nfx, fx = r - w * nfx;
    long long v = 0;
    if (w < W && lane_ok) {
      const size_t o = ((size_t)((a.n0 + w) % a.Lr) * a.NFX + fx) * a.Ntot + tile_gid0 + lane_t;
      v = a.ring_fx[o];
      a.ring_fx[o] = 0;
    }
    fxa[r * TS + lane_t] = v;
  }
  for (int k = tid; k < Wm; k += 256) spn[k] = 0;
  if (tid < ngp * SP) {
    pref[tid] = rq;
    if (tid % SP == 0) gslo[tid / SP] = a.n0 - PL;
  }
  for (int k = tid + 256; k < ngp * SP; k += 256) {  // more buckets than threads
    const int g = k / SP, t = k - g * SP;
    const int pre_pop = a.gproj[tl.gp_begin + g].pre_pop;
    const int s = a.n0 - PL + t;
    pref[k] = (t > 0 && t <= PL) ? (int)a.sb_count[((s - 1) % a.Ls) * a.n_pops + pre_pop] : 0;
    if (t == 0) gslo[g] = a.n0 - PL;
  }
  __syncthreads();
  if (tid < ngp) {  // per-projection inclusive prefix over steps
    int acc = 0;
    for (int t = 0; t < SP; ++t) {
      acc += pref[tid * SP + t];
      pref[tid * SP + t] = acc;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    gtot[0] = 0;
    for (int g = 0; g < ngp; ++g) {
      acc += pref[g * SP + SP - 1];
      gtot[g + 1] = acc;
    }
  }
  __syncthreads();

  // ---- 1. gather: one spike per thread per pass, loads in phases; a
  // segment's edge words come in as aligned 16-byte chunks, three per round
  constexpr int kQ = BS_GATHER_SPIKES, kCH = NARROW ? BS_GATHER_CHUNKS16 : BS_GATHER_CHUNKS;
  constexpr uint32_t kEPC = NARROW ? 8u : 4u;  // edge words per 16-byte chunk
  const int total = gtot[ngp];
  for (int b0 = 0; b0 < total; b0 += kQ * 256) {
    uint32_t si[kQ], e0[kQ], e1[kQ];
    int ss[kQ], sg[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int idx = b0 + q * 256 + tid;
      sg[q] = -1;
      ss[q] = 0;
      si[q] = 0;
      if (idx < total) {
        int g = 0;
        while (gtot[g + 1] <= idx) ++g;
        const int r = idx - gtot[g];
        const int* pg = pref + g * SP;
        int lo = 0, hi = SP - 1;  // pg[lo] <= r < pg[lo + 1]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (pg[mid] <= r) lo = mid; else hi = mid;
        }
        const int s = gslo[g] + lo;
        sg[q] = g;
        ss[q] = s;
        si[q] = a.sb[(long long)(s % a.Ls) * a.Ntot + gi[g].pre_gid0 + (r - pg[lo])];
      }
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      e0[q] = e1[q] = 0;
      if (sg[q] >= 0) {
        const uint32_t* sp = gi[sg[q]].seg + (long long)si[q] * gi[sg[q]].nt1 + tile_in_pop;
        e0[q] = sp[0];
        e1[q] = sp[1];
      }
    }
    const uint4* ep[kQ];
    const long long* qp[kQ];
    int dn[kQ], kc[kQ], cwq[kQ], fxq[kQ], rsq[kQ], baseq[kQ];
    uint32_t incq[kQ], a0[kQ], nchk[kQ], limq[kQ];
    int rounds = 0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int g = sg[q] < 0 ? 0 : sg[q];
      ep[q] = (const uint4*)gi[g].edges;
      qp[q] = gi[g].q;
      kc[q] = gi[g].kind == BS_CH_COUNT;
      cwq[q] = gi[g].cw;
      fxq[q] = gi[g].fx;
      rsq[q] = gi[g].rs;
      incq[q] = 1u << (16 * gi[g].half);
      dn[q] = ss[q] + 1 - a.n0;  // <= 0: the spike's step relative to the window
      // an edge word off = d * RS + lane lands in LDS slot base + off when its
      // arrival offset dn + d is below the horizon H, i.e. off < (H - dn) RS
      baseq[q] = dn[q] * rsq[q] + (kc[q] ? cwq[q] : fxq[q]) * TS;
      limq[q] = (uint32_t)(H - dn[q]) * (uint32_t)rsq[q];
      a0[q] = e0[q] & ~(kEPC - 1u);
      nchk[q] = e1[q] > e0[q] ? (e1[q] - a0[q] + kEPC - 1u) / kEPC : 0u;
      rounds = max(rounds, (int)((nchk[q] + kCH - 1) / kCH));
    }
    for (int rd = 0; rd < rounds; ++rd) {
      uint4 v[kQ][kCH];
#pragma unroll
      for (int q = 0; q < kQ; ++q)
#pragma unroll
        for (int c = 0; c < kCH; ++c) {
          const uint32_t ck = (uint32_t)(rd * kCH + c);
          v[q][c] = ck < nchk[q] ? ep[q][a0[q] / kEPC + ck] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
      for (int q = 0; q < kQ; ++q)
#pragma unroll
        for (int c = 0; c < kCH; ++c) {
          // the chunk's edge words: four 32-bit or eight 16-bit ones
          uint32_t wv[kEPC];
          if constexpr (NARROW) {
            const uint32_t raw[4] = {v[q][c].x, v[q][c].y, v[q][c].z, v[q][c].w};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
              wv[2 * m] = raw[m] & 0xFFFFu;
              wv[2 * m + 1] = raw[m] >> 16;
            }
          } else {
            wv[0] = v[q][c].x;
            wv[1] = v[q][c].y;
            wv[2] = v[q][c].z;
            wv[3] = v[q][c].w;
          }
#pragma unroll
          for (int m = 0; m < (int)kEPC; ++m) {
            const uint32_t e = a0[q] + (uint32_t)((rd * kCH + c) * (int)kEPC + m);
            if (e - e0[q] < e1[q] - e0[q]) {
              const uint32_t off = wv[m];
              if (off < limq[q]) {  // an LDS counter (this window's steps, and later ones when they fit)
                if (kc[q])
                  atomicAdd(&cnt[baseq[q] + (int)off], incq[q]);
                else
                  atomicAdd((unsigned long long*)&fxa[baseq[q] + (int)off], (unsigned long long)qp[q][e]);
              } else {  // beyond the LDS horizon: this tile's slot of the arrival ring
                const int d = (int)(off / (uint32_t)rsq[q]);
                const int ln = (int)(off - (uint32_t)d * (uint32_t)rsq[q]);
                const size_t slot = (size_t)((a.n0 + dn[q] + d) % a.Lr);
                if (kc[q])
                  __hip_atomic_fetch_add(&a.ring_cnt[(slot * a.NCW + cwq[q]) * a.Ntot + tile_gid0 + ln],
                                         incq[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                  __hip_atomic_fetch_add(
                      (unsigned long long*)&a.ring_fx[(slot * a.NFX + fxq[q]) * a.Ntot + tile_gid0 + ln],
                      (unsigned long long)qp[q][e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              }
            }
          }
        }
    }
  }
  __syncthreads();
  // arrivals after this window held in LDS: add them to this tile's ring slots
  for (int k = W * ncw * TS + tid; k < H * ncw * TS; k += 256) {
    const int w = k / (ncw * TS), r = k - w * ncw * TS, cw = r / TS, ln = r - cw * TS;
    const uint32_t v = cnt[k];
    if (v != 0u && ln < tl.count) {
      const size_t o = ((size_t)((a.n0 + w) % a.Lr) * a.NCW + cw) * a.Ntot + tile_gid0 + ln;
      atomicAdd(&a.ring_cnt[o], v);  // result unused: nothing waits for it
    }
  }
  for (int k = W * nfx * TS + tid; k < H * nfx * TS; k += 256) {
    const int w = k / (nfx * TS), r = k - w * nfx * TS, fx = r / TS, ln = r - fx * TS;
    const long long v = fxa[k];
    if (v != 0 && ln < tl.count) {
      const size_t o = ((size_t)((a.n0 + w) % a.Lr) * a.NFX + fx) * a.Ntot + tile_gid0 + ln;
      atomicAdd((unsigned long long*)&a.ring_fx[o], (unsigned long long)v);
    }
  }

  // plastic inputs or outputs: their LDS lies after the window layout
  const bool plp = PLASTIC && (gpop.n_pl > 0 || gpop.n_plpre > 0);
  PlCtx plc{};
  if (plp) {
    const PlLayout PLo = pl_layout(Lo.end, Wm, TS, gpop.n_pl, a.psp, gpop.n_tail, (int)sizeof(T), gpop.pl_cap);
    plc.ppref = (int*)(lds + PLo.ppref);
    plc.poff = (int*)(lds + PLo.poff);
    plc.part = (int*)(lds + PLo.part);
    plc.fm = (unsigned*)(lds + PLo.fm);
    plc.fl = (int*)(lds + PLo.fl);
    plc.tail = lds + PLo.tail;
    plc.rk = (uint32_t*)(lds + PLo.rk);
    plc.rtr = (double*)(lds + PLo.rtr);
    plc.npl = gpop.n_pl;
    plc.nplpre = gpop.n_plpre;
    plc.cap = gpop.pl_cap;
    plc.S = gpop.n_pl * TS;
    plc.tile_in_pop = tile_in_pop;
  }

  // ---- 2. inputs of every (step, lane), channels summed in order, all threads.
  // Channel-outer over the (up to four) items of each thread, so the items'
  // LDS reads of one channel overlap; every item still adds its channels in
  // the summation order.
  {
    const T* tables = (const T*)a.tables;
    __shared__ double schv[32];  // a SCHEDULE channel's value per window step (Wmax <= 32)
    UpdCtx<T> u{a, cnt, fxa, spl, spn, chs, tab, tables, TS, ncw, nfx, nch, tl.start, tid,
                tid < tl.count, (long long)gpop.gid0 + tl.start + tid, inp, nrec, Wm, pc, tl.ncur, schv};
    // (two items per thread in float: fewer registers, more waves per SIMD)
    constexpr int kIT = sizeof(T) == 4 ? 2 : 4;
    const int nit = W * TS;
    // current inputs are summed apart from the delta inputs, starting from
    // I_e + x (x = 0) for iaf_psc_delta and from x = 0 for the latching models
    const T cur0 = gpop.model == BS_IAF_PSC_DELTA ? (T)P[34] + (T)0 : (T)0;
    // chunks of 256 * kIT items, channel-outer inside a chunk
    auto inputs_chunk = [&](const int base0) {
      T acc0[kIT], acc1[kIT], accC[kIT];
      bool ok[kIT];
#pragma unroll
      for (int i = 0; i < kIT; ++i) {
        const int idx = base0 + tid + 256 * i;
        ok[i] = idx < nit && (idx & (TS - 1)) < tl.count;
        acc0[i] = (T)0;
        acc1[i] = (T)0;
        accC[i] = cur0;
      }
      for (int c = 0; c < nch; ++c) {
        const ChReg<T> cc = load_ch<T>(chs[c], tab, tables);
        if (cc.plastic >= 0) continue;  // summed by the update loop, after its arrivals
        if (cc.kind == BS_CH_SCHEDULE) {  // block-uniform: one lookup per window step
          __syncthreads();
          if (tid < W) schv[tid] = sched_at(cc.sched, cc.sval, cc.cdf_len, (long long)a.n0 + tid);
          __syncthreads();
        }
        // one code path per channel kind (the kind is block-uniform), with
        // the staged tables addressed in LDS
        T v[kIT];
        const long long loff = chs[c].lds_off;
        if (cc.kind == BS_CH_COUNT) {
          const T* tl_ = (const T*)(tab + (loff >= 0 ? loff : 0));
          const uint32_t nl = loff >= 0 ? (uint32_t)cc.tab_n : 0u;
#pragma unroll
          for (int i = 0; i < kIT; ++i) {
            const int idx = base0 + tid + 256 * i, w = idx >> lgTS, l = idx & (TS - 1);
            v[i] = (T)0;
            if (ok[i]) {
              uint32_t k = (cnt[(w * ncw + cc.cw) * TS + l] >> (16 * cc.half)) & 0xFFFFu;
              if (k > (uint32_t)cc.max_count) {  // cannot happen for consistent edges
                atomicOr(a.flags, FLAG_COUNT);
                k = (uint32_t)cc.max_count;
              }
              v[i] = k < nl ? tl_[k] : cc.tabg[k];
            }
          }
        } else if (cc.kind == BS_CH_POISSON && !cc.pre && loff >= 0) {
          const double* cdf = (const double*)(tab + loff);
          const unsigned short* guide = (const unsigned short*)(cdf + cc.cdf_len);
#pragma unroll
          for (int i = 0; i < kIT; ++i) {
            const int idx = base0 + tid + 256 * i, w = idx >> lgTS, l = idx & (TS - 1);
            v[i] = ok[i] ? pois_value<T, false>(cc, cdf, guide, tl.start + l, a.n0 + w, nullptr) : (T)0;
          }
        } else {
#pragma unroll
          for (int i = 0; i < kIT; ++i) {
            const int idx = base0 + tid + 256 * i, w = idx >> lgTS, l = idx & (TS - 1);
            v[i] = ok[i] ? chan_value<T>(u, cc, c, w, a.n0 + w, l) : (T)0;
          }
        }
        if (cc.tail >= 0) {  // summed after a plastic channel: kept for the update loop
          T* tw = (T*)plc.tail;
#pragma unroll
          for (int i = 0; i < kIT; ++i) {
            const int idx = base0 + tid + 256 * i, w = idx >> lgTS, l = idx & (TS - 1);
            if (ok[i]) tw[(cc.tail * Wm + w) * TS + l] = v[i];
          }
        } else if (cc.kind == BS_CH_CURRENT || cc.kind == BS_CH_AC || cc.kind == BS_CH_SCHEDULE) {
#pragma unroll
          for (int i = 0; i < kIT; ++i) accC[i] = accC[i] + v[i];
        } else if (cc.receptor == 0) {
#pragma unroll
          for (int i = 0; i < kIT; ++i) acc0[i] = acc0[i] + v[i];
        } else {
#pragma unroll
          for (int i = 0; i < kIT; ++i) acc1[i] = acc1[i] + v[i];
        }
      }
#pragma unroll
      for (int i = 0; i < kIT; ++i) {
        const int idx = base0 + tid + 256 * i, w = idx >> lgTS, l = idx & (TS - 1);
        if (ok[i]) {
          inp[w * TS + l] = acc0[i];
          if (nrec > 1) inp[(Wm + w) * TS + l] = acc1[i];
          if (tl.ncur) inp[(nrec * Wm + w) * TS + l] = accC[i];
        }
      }
    };
    for (int base0 = 0; base0 < nit; base0 += 256 * kIT) inputs_chunk(base0);
  }
  __syncthreads();

  // ---- 3. update: the W-step loop of each neuron (populations with plastic
  // projections: all threads, for the block-wide arrival lists)
  if (plp || tid < TS) {
    const T* tables = (const T*)a.tables;
    UpdCtx<T> u{a, cnt, fxa, spl, spn, chs, tab, tables, TS, ncw, nfx, nch, tl.start, tid,
                tid < tl.count, (long long)gpop.gid0 + tl.start + tid, inp, nrec, Wm, pc, tl.ncur, nullptr};
    // each model with and without current inputs (a compile-time branch)
#define BS_UPDATE(CUR)                                                                             \
    switch (gpop.model) {                                                                         \
      case BS_IAF_PSC_DELTA: {                                                                    \
        const DeltaP<T> mp = load_delta<T>(P);                                                    \
        const auto f = [&](T* s, int& r, T a0, T, T cur) { return step_delta<T, CUR>(mp, s, r, a0, cur); }; \
        StepLane<T, decltype(f)> m(gpop.n_state, f);                                              \
        run_update<T, CUR>(u, plp, plc, gpop, m);                                                 \
      } break;                                                                                    \
      case BS_IAF_PSC_EXP: {                                                                      \
        const ExpP<T> mp = load_exp<T>(P);                                                        \
        const auto f = [&](T* s, int& r, T a0, T, T cur) { return step_exp<T>(mp, s, r, a0, cur); }; \
        StepLane<T, decltype(f)> m(gpop.n_state, f);                                              \
        run_update<T, CUR>(u, plp, plc, gpop, m);                                                 \
      } break;                                                                                    \
      case BS_IAF_PSC_ALPHA: {                                                                    \
        const AlphaP<T> mp = load_alpha<T>(P);                                                    \
        const auto f = [&](T* s, int& r, T a0, T, T cur) { return step_alpha<T>(mp, s, r, a0, cur); }; \
        StepLane<T, decltype(f)> m(gpop.n_state, f);                                              \
        run_update<T, CUR>(u, plp, plc, gpop, m);                                                 \
      } break;                                                                                    \
      case BS_AEIF_COND_ALPHA: {                                                                  \
        if constexpr (std::is_same<T, double>::value) {                                           \
          AeifRun m(P);                                                                           \
          run_update<T, CUR>(u, plp, plc, gpop, m);                                               \
          if (m.unstable) atomicOr(a.flags, FLAG_UNSTABLE);                                       \
        }                                                                                         \
      } break;                                                                                    \
    }
    if (tl.ncur) {
      BS_UPDATE(true)
    } else {
      BS_UPDATE(false)
    }
#undef BS_UPDATE
  }
  __syncthreads();

  // ---- 4. emit: one bucket and one record reservation per step
  if (tid < W && spn[tid] > 0) {
    const int step = a.n0 + tid;
    base[2 * tid] = atomicAdd(&a.sb_count[(step % a.Ls) * a.n_pops + tl.pop], (unsigned)spn[tid]);
    base[2 * tid + 1] = atomicAdd(a.rec_count, (unsigned long long)spn[tid]);
  }
  __syncthreads();
  for (int w = 0; w < W; ++w) {
    const int m = spn[w];
    for (int k = tid; k < m; k += 256) {
      const int jj = tl.start + (int)spl[w * TS + k];
      const int step = a.n0 + w;
      a.sb[(long long)(step % a.Ls) * a.Ntot + gpop.gid0 + (long long)base[2 * w] + k] = (uint32_t)jj;
      const unsigned long long idx = base[2 * w + 1] + (unsigned long long)k;
      if ((long long)idx < a.rec_cap)
        a.rec[idx] = ((unsigned long long)(step - a.vm_step0) << a.gid_bits) |
                     (unsigned long long)(gpop.gid0 + jj);
      else
        atomicOr(a.flags, FLAG_OVERFLOW);
    }
  }
}

template <typename T, bool PLASTIC, bool NARROW>
__global__ __launch_bounds__(256) void k_window(WinArgs a) {
  window_body<T, PLASTIC, NARROW>(a);
}
// float networks without plastic projections: six waves per SIMD (the body
// fits in 80 VGPRs with two inputs-phase items per thread; the ensemble's
// latency-bound blocks gain from every extra wave: 4 -> 5 -> 6 waves took a
// 256-network window from 648 to 589 to 533 us, 7 waves with one item per
// thread 579 us)
template <>
__global__ __launch_bounds__(256, 6) void k_window<float, false, false>(WinArgs a) {
  window_body<float, false, false>(a);
}
template <>
__global__ __launch_bounds__(256, 6) void k_window<float, false, true>(WinArgs a) {
  window_body<float, false, true>(a);
}

// ============================================================== host side
struct bs_net {
  int dtype = BS_F64;
  size_t tsize = 8;
  double dt = 0.1;
  hipStream_t stream = nullptr;
  int device = 0;
  // layout
  long long Ntot = 0;
  int W = 1, d_min = 0, d_max = 0, Ls = 1, max_span = 1, TS = 256;
  int psp = 1;            // plastic gather: source steps per projection and window, at most
  bool narrow = false;    // static edge words are 16-bit
  size_t lds_bytes = 0;
  size_t lds_budget = 0;  // dynamic LDS the window kernel may use
  std::vector<DevPop> pops;
  std::vector<DevChan> chans;
  std::vector<Tile> tiles;
  std::vector<int> pop_record;
  std::vector<int> pop_remote;  // 1: spikes injected by bs_net_inject_spikes
  bool all_recorded = true;  // every population keeps its spikes for the caller
  // device buffers
  Tile* d_tiles = nullptr;
  DevPop* d_pops = nullptr;
  DevChan* d_chans = nullptr;
  GProj* d_gproj = nullptr;
  void* d_tables = nullptr;
  void* d_state = nullptr;
  int* d_ref = nullptr;
  uint32_t* d_sb = nullptr;
  uint32_t* d_sb_count = nullptr;
  // Spike records are double-buffered: chunk c writes buffer c & 1 while the
  // host sorts, copies and decodes chunk c - 1 (second stream, pinned memory).
  unsigned long long* d_rec[2] = {nullptr, nullptr};  // record keys of a chunk, emit order
  unsigned long long* d_rec_sorted = nullptr;
  void* d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int gid_bits = 1, step_bits = 1;
  unsigned long long* d_rec_count = nullptr;  // [2], one per record buffer
  long long rec_cap = 0;
  int rec_par = 0;                             // record buffer the next launches write
  hipStream_t stream2 = nullptr;               // record sorts and their copies
  hipEvent_t ev_chunk[2] = {nullptr, nullptr};
  hipEvent_t ev_join[2] = {nullptr, nullptr};  // bs_net_run_device: caller stream -> ours -> caller's
  int* d_tap_col = nullptr;                    // device column of each voltmeter tap, caller order
  // record_weight taps: plastic projections whose weights are copied after
  // every step into a device trajectory, then into host memory per chunk
  struct WTap {
    int q = 0;                                // stdp index
    long long E = 0;
    long long* d_ref = nullptr;               // reference (CSR) position -> stored position
    double* d_traj = nullptr;                 // [wtap_rows][E]
    std::vector<double> host;                 // [n_steps][E] of the last run
  };
  std::vector<WTap> wtaps;
  long long wtap_rows = 0;                    // device trajectory rows (steps per chunk)
  long long wtap_steps = 0;                   // steps in the host trajectories
  // ac_generator channels: {omega, phi, dt, t_min_step, t_max_step}; a run
  // whose sine arguments would leave the sine's exact range is refused
  std::vector<std::array<double, 5>> ac_args;
  // pinned host memory of the run loop: per-buffer count, flags, chunk base
  // step and previous window length, then the sorted records and V_m rows
  unsigned long long* h_cnt = nullptr;
  int* h_flags = nullptr;
  long long* h_base_p = nullptr;
  int* h_prev_p = nullptr;
  unsigned long long* h_rec = nullptr;
  size_t h_rec_cap = 0;
  // record buffers outgrown during a run: hipHostFree waits for the whole
  // device, so they are freed after the run's last chunk, not between chunks
  std::vector<void*> h_rec_old;
  double* h_vm_p = nullptr;                    // [2][vm_cap_steps * n_vm_dev]
  size_t h_vm_cap = 0;
  int* d_flags = nullptr;
  int* d_vm_col = nullptr;
  int n_vm_dev = 0;
  double* d_vm = nullptr;
  long long vm_cap_steps = 0;
  // plastic projections
  std::vector<StdpDev> stdp;         // device pointers inside
  StdpDev* d_stdp = nullptr;
  unsigned char* d_spk_hist = nullptr;
  std::vector<int> proj_stdp;        // descriptor projection -> stdp index (-1: static)
  std::vector<std::vector<double>> w0;        // initial weights, (pre, delay) order
  std::vector<std::vector<long long>> w_ref;  // reference CSR position -> stored position
  long long* d_base = nullptr;     // first step of the chunk a graph replay advances
  // arrival ring (events landing after their delivery window), per tile slice
  uint32_t* d_ring_cnt = nullptr;
  long long* d_ring_fx = nullptr;
  unsigned char* d_blob = nullptr;          // per-population blobs for k_window
  std::vector<unsigned char> h_blob;
  std::vector<std::vector<double>> h_cdf;   // Poisson tables per channel (host copies)
  std::vector<std::vector<unsigned short>> h_guide;
  int Lr = 1, NCW = 1, NFX = 0;
  int H = 1;                       // arrival offsets held in LDS by k_window
  int* d_prevlen = nullptr;        // graph replays: length of the window before the chunk
  int last_win = 0;                // length of the last window run (0 after reset)
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // one full chunk of window launches per record buffer
  // Poisson producer blocks (when the tiles leave CUs idle)
  std::vector<int> pop_pc;         // population has a drawn-ahead Poisson channel
  std::vector<PJob> pjobs;
  PJob* d_pjobs = nullptr;
  unsigned short* d_pcnt = nullptr;
  long long pjob_n = 0;
  int n_prod = 0;
  std::vector<void*> owned;  // per-projection gather arrays, generator maps
  std::vector<int32_t*> drive;     // per channel: counts injected by bs_net_set_drive (owned), or null
  std::vector<long long> pop_blob;  // per population: offset of its blob in d_blob (-1: none)
  // host copies
  std::vector<double> state0;  // [BS_NSTATE][Ntot]
  std::vector<int> ref0;
  std::vector<long long> outdeg;  // static out-degree per gid (event accounting)
  std::vector<int> gid_pop;       // population of each gid
  std::vector<int> vm_tap_col;    // per tap (in caller order) -> device column
  long long step = 0;
  // decoded spikes of the last run, one block per chunk; blocks keep their
  // capacity across runs so a steady run decodes into memory already touched
  struct SpikeBlock {
    std::vector<int32_t> gid, step;
    size_t n = 0;
  };
  std::vector<SpikeBlock> out_blocks;
  size_t n_out_blocks = 0;
  long long n_out = 0;
  // timing / accounting of the last run
  bool timing = false;
  double t_deliver = 0, t_update = 0, t_total = 0;
  long long n_deliver = 0, n_update = 0;
  double b_deliver = 0, b_update = 0;
  long long n_events = 0;
};

// Public entry points keep the C linkage of their declarations in the header.
int bs_abi_version(void) { return BS_ABI_VERSION; }
#ifndef BS_SOURCE_HASH
#define BS_SOURCE_HASH "unknown"
#endif
const char* bs_source_hash(void) { return BS_SOURCE_HASH; }
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
  F(net->d_tiles); F(net->d_pops); F(net->d_chans); F(net->d_gproj); F(net->d_tables);
  F(net->d_state); F(net->d_ref); F(net->d_sb); F(net->d_sb_count);
  F(net->d_rec[0]); F(net->d_rec[1]); F(net->d_rec_sorted); F(net->d_sort_tmp);
  F(net->d_rec_count); F(net->d_flags); F(net->d_vm_col); F(net->d_vm); F(net->d_tap_col);
  for (void* p : net->owned) F(p);
  for (int32_t* p : net->drive) F(p);
  for (auto& t : net->wtaps) {
    F(t.d_ref);
    F(t.d_traj);
  }
  F(net->d_base); F(net->d_blob); F(net->d_ring_cnt); F(net->d_ring_fx); F(net->d_prevlen);
  F(net->d_stdp);
  F(net->d_spk_hist);
  F(net->d_pjobs);
  F(net->d_pcnt);
  for (int k = 0; k < 2; ++k) {
    if (net->graph[k]) (void)hipGraphExecDestroy(net->graph[k]);
    if (net->ev_chunk[k]) (void)hipEventDestroy(net->ev_chunk[k]);
    if (net->ev_join[k]) (void)hipEventDestroy(net->ev_join[k]);
  }
  auto FH = [](void* p) {
    if (p) (void)hipHostFree(p);
  };
  FH(net->h_cnt); FH(net->h_flags); FH(net->h_base_p); FH(net->h_prev_p); FH(net->h_rec); FH(net->h_vm_p);
  for (void* p : net->h_rec_old) FH(p);
  if (net->stream2) (void)hipStreamDestroy(net->stream2);
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
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(X));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return BS_OK;
}

template <typename X>
static int upload(bs_net* net, X** d, const X* h, size_t n) {
  int rc = dmalloc(d, n);
  if (rc) return rc;
  net->owned.push_back((void*)*d);
  if (n) HIPCHK(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
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
static const size_t kLdsBudget = 64 * 1024;
static const int kStageCount = 256;  // count-table entries staged in LDS per channel

// CDF table of Poisson(mu), mu = lam / parts, parts = max(1, ceil(lam / 32)):
// pmf_0 = exp(-mu), pmf_k = pmf_{k-1} * mu / k, cdf_k = cdf_{k-1} + pmf_k,
// until k > mu and 1 - cdf_k < 1e-15 (at most 512 entries); guide[g] is the
// first k with cdf_k >= g / 64. poisson_inv samples by inversion of it.
static void poisson_table(double lam, std::vector<double>& cdf, std::vector<unsigned short>& guide,
                          int& parts) {
  parts = std::max(1, (int)std::ceil(lam / 32.0));
  const double mu = lam / (double)parts;
  double pmf = std::exp(-mu);
  cdf.assign(1, pmf);
  for (int k = 1; k < 512; ++k) {
    pmf = pmf * mu / (double)k;
    const double c = cdf.back() + pmf;
    cdf.push_back(c);
    if ((double)k > mu && 1.0 - c < 1e-15) break;
  }
  const int len = (int)cdf.size();
  guide.assign(64, 0);
  for (int g = 0; g < 64; ++g) {
    const double t = (double)g / 64.0;
    int k = 0;
    while (k < len - 1 && cdf[k] < t) ++k;
    guide[g] = (unsigned short)k;
  }
}

// ---- gather segments of a projection whose edges are already in device
// memory: the layout of the host counting sort (segments in (pre, tile)
// row-major order, each holding its edges in edge order), from a stable radix
// sort of the edge indices by (pre, tile).
__global__ void k_seg_keys(const int32_t* __restrict__ pre, const int32_t* __restrict__ post,
                           const int32_t* __restrict__ delay, int delay_steps, long long E, int npre, int npost,
                           int TS, int nt, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                           int* __restrict__ stats) {
  // grid-stride; with per-edge delays one min/max atomic pair per block
  int dmn = 0x7FFFFFFF, dmx = -1;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
    const int a = pre[e], b = post[e];
    const int dd = delay ? delay[e] : delay_steps;
    uint32_t key = 0;
    if (a < 0 || a >= npre || b < 0 || b >= npost) {
      atomicOr(&stats[0], 1);
    } else if (dd < 0 || dd >= (1 << 24)) {
      atomicOr(&stats[0], 2);
    } else {
      key = (uint32_t)a * (uint32_t)nt + (uint32_t)(b / TS);
      dmn = min(dmn, dd);
      dmx = max(dmx, dd);
    }
    keys[e] = key;
    vals[e] = (uint32_t)e;
  }
  if (!delay) return;  // a homogeneous delay is known on the host
  __shared__ int red[2][4];
  for (int o = 32; o > 0; o >>= 1) {
    dmn = min(dmn, __shfl_xor(dmn, o, 64));
    dmx = max(dmx, __shfl_xor(dmx, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = dmn;
    red[1][threadIdx.x >> 6] = dmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int mn = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
    const int mx = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
    if (mx >= 0) {
      atomicMin(&stats[1], mn);
      atomicMax(&stats[2], mx);
    }
  }
}

// seg[a * (nt + 1) + t] = first sorted position with key >= a * nt + t; the
// entry t = nt is the row end.
__global__ void k_seg_bounds(const uint32_t* __restrict__ keys, long long E, int npre, int nt,
                             uint32_t* __restrict__ seg, int* __restrict__ deg) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)npre * (nt + 1);
  if (idx >= n) return;
  const long long a = idx / (nt + 1), t = idx - a * (nt + 1);
  const unsigned long long target = (unsigned long long)a * nt + t;
  long long lo = 0, hi = E;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((unsigned long long)keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  seg[idx] = (uint32_t)lo;
  if (t == nt) {  // the row is complete once its end is known: one thread per row
    long long lo0 = 0, hi0 = E;
    const unsigned long long t0 = (unsigned long long)a * nt;
    while (lo0 < hi0) {
      const long long mid = (lo0 + hi0) >> 1;
      if ((unsigned long long)keys[mid] < t0) lo0 = mid + 1;
      else hi0 = mid;
    }
    deg[a] = (int)(lo - lo0);
  }
}

__global__ void k_seg_edges(const uint32_t* __restrict__ order, long long E, const int32_t* __restrict__ post,
                            const int32_t* __restrict__ delay, int delay_steps, const double* __restrict__ w,
                            int F, int TS, int RS, uint32_t* __restrict__ ew, long long* __restrict__ eq) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E + 4) return;
  if (k >= E) {  // the gather reads whole 16-byte chunks past a segment's end
    ew[k] = 0u;
    return;
  }
  const uint32_t e = order[k];
  const int dd = delay ? delay[e] : delay_steps;
  ew[k] = (uint32_t)dd * (uint32_t)RS + (uint32_t)(post[e] % TS);
  if (eq) eq[k] = (long long)llrint(ldexp(w[e], F));
}

// 16-bit copy of a projection's edge words (each below 2^16), zero-padded by
// eight words: the gather reads whole 16-byte chunks past a segment's end.
__global__ void k_narrow_edges(const uint32_t* __restrict__ in, long long E, uint16_t* __restrict__ out) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < E + 8;
       k += (long long)gridDim.x * blockDim.x)
    out[k] = k < E ? (uint16_t)in[k] : (uint16_t)0;
}

// Scratch of the device segment build, grown to the largest projection.
struct SegScratch {
  uint32_t* kv[4] = {nullptr, nullptr, nullptr, nullptr};  // keys in/out, values in/out
  long long cap = 0;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int* stats = nullptr;
  int* deg = nullptr;
  long long deg_cap = 0;
  ~SegScratch() {
    for (uint32_t* p : kv) (void)hipFree(p);
    (void)hipFree(tmp);
    (void)hipFree(stats);
    (void)hipFree(deg);
  }
};

// Builds G.seg / G.edges / G.q of a static projection with device-resident
// edges on net->stream; adds the out-degrees to net->outdeg and returns the
// projection's delay range. Messages match the host build's.
static int build_segments_device(bs_net* net, const bs_projection& pj, int kind, int F, int RS, int npre,
                                 int npost, long long g0, SegScratch& S, GProj& G, int& pdmin, int& pdmax) {
  const int TS = net->TS;
  const long long E = pj.n_edges;
  const int nt = (npost + TS - 1) / TS;
  if ((unsigned long long)npre * (unsigned long long)nt >= (1ull << 32))
    return fail(BS_EINVAL, "device edges: more than 2^32 (pre, tile) segments in one projection");
  if (E >= (1ll << 31)) return fail(BS_EINVAL, "device edges: more than 2^31 - 1 edges in one projection");
  if (E > 0 && (!pj.pre || !pj.post)) return fail(BS_EINVAL, "device edges: null edge array");
  hipStream_t st = net->stream;
  if (E > S.cap) {
    for (uint32_t*& p : S.kv) {
      (void)hipFree(p);
      p = nullptr;
    }
    S.cap = 0;
    for (uint32_t*& p : S.kv) HIPCHK(hipMalloc(&p, sizeof(uint32_t) * (size_t)E));
    S.cap = E;
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, S.kv[0], S.kv[1], S.kv[2], S.kv[3], (int)E, 0, 32, st));
    if (tb > S.tmp_bytes) {
      (void)hipFree(S.tmp);
      S.tmp = nullptr;
      S.tmp_bytes = 0;
      HIPCHK(hipMalloc(&S.tmp, tb));
      S.tmp_bytes = tb;
    }
  }
  if (!S.stats) HIPCHK(hipMalloc(&S.stats, 4 * sizeof(int)));
  if (npre > S.deg_cap) {
    (void)hipFree(S.deg);
    S.deg = nullptr;
    HIPCHK(hipMalloc(&S.deg, sizeof(int) * (size_t)std::max(1, npre)));
    S.deg_cap = npre;
  }
  const int stats0[4] = {0, 0x7FFFFFFF, -1, 0};
  HIPCHK(hipMemcpyAsync(S.stats, stats0, sizeof(stats0), hipMemcpyHostToDevice, st));
  int bits = 1;
  while (bits < 32 && (1ull << bits) < (unsigned long long)npre * nt) ++bits;
  if (E > 0) {
    const long long nb = std::min<long long>((E + 255) / 256, 8192);
    hipLaunchKernelGGL(k_seg_keys, dim3((unsigned)nb), dim3(256), 0, st, pj.pre, pj.post, pj.delay,
                       pj.delay_steps, E, npre, npost, TS, nt, S.kv[0], S.kv[2], S.stats);
    HIPCHK(hipGetLastError());
    size_t tb = S.tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(S.tmp, tb, S.kv[0], S.kv[1], S.kv[2], S.kv[3], (int)E, 0, bits, st));
  }
  uint32_t *dseg = nullptr, *dew = nullptr;
  long long* deq = nullptr;
  const long long nseg = (long long)npre * (nt + 1);
  int rc = dmalloc(&dseg, (size_t)std::max<long long>(1, nseg));
  if (rc) return rc;
  net->owned.push_back(dseg);
  rc = dmalloc(&dew, (size_t)E + 4);
  if (rc) return rc;
  net->owned.push_back(dew);
  if (kind == BS_CH_FIXED) {
    rc = dmalloc(&deq, (size_t)std::max<long long>(1, E));
    if (rc) return rc;
    net->owned.push_back(deq);
  }
  if (nseg > 0) {
    hipLaunchKernelGGL(k_seg_bounds, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, st, S.kv[1], E, npre, nt,
                       dseg, S.deg);
    HIPCHK(hipGetLastError());
  }
  int stats[4];
  std::vector<int> deg((size_t)npre);
  HIPCHK(hipMemcpyAsync(stats, S.stats, sizeof(stats), hipMemcpyDeviceToHost, st));
  if (npre > 0) HIPCHK(hipMemcpyAsync(deg.data(), S.deg, sizeof(int) * (size_t)npre, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stats[0] & 1) return fail(BS_EINVAL, "edge index out of range");
  if (stats[0] & 2) return fail(BS_EINVAL, "delay must be in [0, 2^24) steps");
  if (E == 0) {
    pdmin = pdmax = 0;
  } else if (!pj.delay) {
    pdmin = pdmax = pj.delay_steps;
  } else {
    pdmin = stats[1];
    pdmax = stats[2];
  }
  // edge words delay * RS + lane must fit 32 bits
  if ((unsigned long long)std::max(0, pdmax) * (unsigned long long)RS + (unsigned long long)TS > 0xFFFFFFFFull)
    return fail(BS_EINVAL, "delay too long for this population's arrival counters (delay * stride >= 2^32)");
  hipLaunchKernelGGL(k_seg_edges, dim3((unsigned)((E + 4 + 255) / 256)), dim3(256), 0, st, S.kv[3], E, pj.post,
                     pj.delay, pj.delay_steps, kind == BS_CH_FIXED ? pj.weight : nullptr, F, TS, RS, dew, deq);
  HIPCHK(hipGetLastError());
  for (int a = 0; a < npre; ++a) net->outdeg[g0 + a] += deg[a];
  G.seg = dseg;
  G.edges = dew;
  G.q = deq;
  G.nt = nt;
  return BS_OK;
}

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
  HCHK(hipStreamCreateWithFlags(&net->stream2, hipStreamNonBlocking));
  for (int k = 0; k < 2; ++k) HCHK(hipEventCreateWithFlags(&net->ev_chunk[k], hipEventDisableTiming));
  for (int k = 0; k < 2; ++k) HCHK(hipEventCreateWithFlags(&net->ev_join[k], hipEventDisableTiming));
  HCHK(hipHostMalloc((void**)&net->h_cnt, 2 * sizeof(unsigned long long), hipHostMallocDefault));
  HCHK(hipHostMalloc((void**)&net->h_flags, 2 * sizeof(int), hipHostMallocDefault));
  HCHK(hipHostMalloc((void**)&net->h_base_p, 2 * sizeof(long long), hipHostMallocDefault));
  HCHK(hipHostMalloc((void**)&net->h_prev_p, 2 * sizeof(int), hipHostMallocDefault));

  // ---- tile size: the largest of 256/128/64 neurons that still gives >= 1024
  // tiles (four per CU), so small networks spread over the whole GPU. With
  // aeif_cond_alpha the update (adaptive RKF45) dominates the window, and all
  // four waves of a block take part in it only at 256 neurons per tile, so
  // such networks keep the largest tiles that still give >= 128 blocks.
  int TS = 256;
  {
    auto ntiles = [&](int ts) {
      long long t = 0;
      for (int p = 0; p < d->n_pops; ++p) t += (d->pops[p].n + ts - 1) / ts;
      return t;
    };
    bool rkf = false;
    for (int p = 0; p < d->n_pops; ++p) rkf = rkf || d->pops[p].model == BS_AEIF_COND_ALPHA;
    const long long min_tiles = rkf ? 128 : 1024;
    while (TS > 64 && ntiles(TS) < min_tiles) TS /= 2;
    // fewer 64-neuron tiles than CUs: the window is one partial round bound by
    // each block's chain of dependent phases, and 32-neuron tiles halve the
    // gather and inputs work of that chain (Brunel-11250: 24.7 -> 23.5 us)
    int n_cu = 0;
    HCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, net->device));
    if (!rkf && TS == 64 && ntiles(64) < n_cu) TS = 32;
    if (d->tile_size != 0) {
      if (d->tile_size != 32 && d->tile_size != 64 && d->tile_size != 128 && d->tile_size != 256)
        BAD("bs_net_create: tile_size must be 0, 32, 64, 128 or 256");
      TS = d->tile_size;
    }
  }
  net->TS = TS;

  // ---- populations, channels, tiles
  long long gid = 0;
  net->pops.resize(d->n_pops);
  net->chans.resize(d->n_channels);
  net->h_cdf.assign(d->n_channels, {});
  net->h_guide.assign(d->n_channels, {});
  for (int c = 0; c < d->n_channels; ++c) {
    memset(&net->chans[c], 0, sizeof(DevChan));
    net->chans[c].plastic = -1;
    net->chans[c].tail = -1;
  }
  std::vector<int> chan_pop(d->n_channels, -1);
  // channels fed by plastic projections take no fixed-point arrival slot:
  // their sums are formed in the update loop
  std::vector<char> chan_plastic(d->n_channels, 0);
  for (int q = 0; q < d->n_projs; ++q)
    if (d->projs[q].plastic && d->projs[q].channel >= 0 && d->projs[q].channel < d->n_channels)
      chan_plastic[d->projs[q].channel] = 1;
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
    if (sp.remote && (sp.ch_count != 0 || sp.n_vm != 0))
      BAD("a remote population has no input channels and no voltmeter taps");
    dp.ch_begin = sp.ch_begin;
    dp.ch_count = sp.ch_count;
    if (sp.ch_begin < 0 || sp.ch_count < 0 || sp.ch_begin + sp.ch_count > d->n_channels)
      BAD("population channel range out of bounds");
    int ncount = 0, nfx = 0;
    for (int c = sp.ch_begin; c < sp.ch_begin + sp.ch_count; ++c) {
      const bs_channel& ch = d->channels[c];
      DevChan& dc = net->chans[c];
      if (chan_pop[c] >= 0) BAD("a channel belongs to two populations");
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
        dc.fx = chan_plastic[c] ? -1 : nfx++;
        dc.shift = ch.fixed_shift;
        if (ch.fixed_shift < 0 || ch.fixed_shift > 1000) BAD("bad fixed_shift");
      } else if (ch.kind == BS_CH_POISSON) {
        if (!(ch.lam >= 0)) BAD("Poisson lambda must be >= 0");
        std::vector<double> cdf;
        std::vector<unsigned short> guide;
        poisson_table(ch.lam, cdf, guide, dc.parts);
        dc.cdf_len = (int)cdf.size();
        net->h_cdf[c] = cdf;
        net->h_guide[c] = guide;
        double* dcdf = nullptr;
        unsigned short* dgd = nullptr;
        CHK(upload(net, &dcdf, cdf.data(), cdf.size()));
        CHK(upload(net, &dgd, guide.data(), guide.size()));
        dc.cdf = dcdf;
        dc.guide = dgd;
        dc.tmin = ch.t_min_step;
        dc.tmax = ch.t_max_step;
        dc.seed = ch.seed;
        dc.delay = ch.delay_steps;
        dc.frac = ch.delay_frac;
        if (ch.delay_steps < 0) BAD("Poisson delay must be >= 0");
        if (!(ch.delay_frac >= 0.0 && ch.delay_frac < 1.0)) BAD("Poisson delay_frac must be in [0, 1)");
        if (ch.lam == 0.0) dc.tmax = -1;  // rate <= 0: all-zero output
        dc.n_gen = sp.n;
        if (ch.gen_index) {
          int* g = nullptr;
          CHK(upload(net, &g, ch.gen_index, (size_t)sp.n));
          dc.gen_index = g;
          int mx = -1;
          for (int j = 0; j < sp.n; ++j) mx = std::max(mx, ch.gen_index[j]);
          dc.n_gen = mx + 1;
        }
      } else if (ch.kind == BS_CH_CURRENT) {
        if (!(ch.weight == ch.weight)) BAD("current channel value must not be NaN");
        if (ch.receptor != 0) BAD("current channels feed the current-input seam (receptor 0)");
        dc.weight = ch.weight;
        dc.tmin = ch.t_min_step;
        dc.tmax = ch.t_max_step;
        if (ch.gen_index) {
          int* g = nullptr;
          CHK(upload(net, &g, ch.gen_index, (size_t)sp.n));
          dc.gen_index = g;
        }
        ++dp.n_cur;
      } else if (ch.kind == BS_CH_SCHEDULE) {
        if (ch.receptor != 0) BAD("current channels feed the current-input seam (receptor 0)");
        if (ch.sched_len < 1 || !ch.sched || !ch.sched_val) BAD("schedule channel needs sched_len >= 1 plateaus");
        for (int k = 0; k < ch.sched_len; ++k) {
          if (!(ch.sched[2 * k] < ch.sched[2 * k + 1])) BAD("schedule plateaus must have start < stop");
          if (k > 0 && ch.sched[2 * k] < ch.sched[2 * k - 1]) BAD("schedule plateaus must be sorted and disjoint");
          if (!(ch.sched_val[k] == ch.sched_val[k])) BAD("schedule values must not be NaN");
        }
        long long* dse = nullptr;
        double* dsv = nullptr;
        CHK(upload(net, &dse, (const long long*)ch.sched, 2 * (size_t)ch.sched_len));
        CHK(upload(net, &dsv, ch.sched_val, (size_t)ch.sched_len));
        dc.sched = dse;
        dc.sval = dsv;
        dc.cdf_len = ch.sched_len;
        dc.tmin = ch.sched[0];
        dc.tmax = ch.sched[2 * ch.sched_len - 1];
        if (ch.gen_index) {
          int* g = nullptr;
          CHK(upload(net, &g, ch.gen_index, (size_t)sp.n));
          dc.gen_index = g;
        }
        ++dp.n_cur;
      } else if (ch.kind == BS_CH_AC) {
        const double pr[6] = {ch.ac[0], ch.ac[1], ch.ac[2], ch.ac[3], ch.weight, d->dt_ms};
        for (double x : pr)
          if (!std::isfinite(x)) BAD("ac channel parameters must be finite");
        if (ch.receptor != 0) BAD("current channels feed the current-input seam (receptor 0)");
        double* dpr = nullptr;
        CHK(upload(net, &dpr, pr, 6));
        dc.cdf = dpr;
        dc.weight = ch.weight;
        dc.tmin = ch.t_min_step;
        dc.tmax = ch.t_max_step;
        net->ac_args.push_back({ch.ac[1], ch.ac[2], d->dt_ms, (double)ch.t_min_step, (double)ch.t_max_step});
        if (ch.gen_index) {
          int* g = nullptr;
          CHK(upload(net, &g, ch.gen_index, (size_t)sp.n));
          dc.gen_index = g;
        }
        ++dp.n_cur;
      } else {
        BAD("unknown channel kind");
      }
    }
    dp.n_cw = (ncount + 1) / 2;
    dp.n_fx = nfx;
    dp.tile0 = (int)net->tiles.size();
    if (!sp.remote)  // a remote population's neurons are updated by their own shard
      for (int s0 = 0; s0 < sp.n; s0 += TS)
        net->tiles.push_back(Tile{p, s0, std::min(TS, sp.n - s0), 0});
    gid += sp.n;
    const int rec = sp.remote ? 0 : sp.record_spikes;
    net->pop_record.push_back(rec);
    net->pop_remote.push_back(sp.remote ? 1 : 0);
    if (!rec) net->all_recorded = false;
  }
  net->Ntot = gid;
  if (gid >= (1ll << 31)) BAD("more than 2^31 neurons");
  net->gid_pop.assign(gid, 0);
  for (int p = 0; p < d->n_pops; ++p)
    for (int i = 0; i < net->pops[p].n; ++i) net->gid_pop[net->pops[p].gid0 + i] = p;

  // ---- static projections -> per-(pre, post tile) gather segments
  net->outdeg.assign(gid, 0);
  std::vector<GProj> gp_h;
  std::vector<std::vector<int>> pop_gp(d->n_pops);
  std::vector<int> chan_proj(d->n_channels, -1);
  int dmin = 1 << 30, dmax = 0;
  // Static projections: the per-(pre, tile) segments are a counting sort of
  // each projection's edges, built for all projections in parallel on host
  // threads (the order inside a segment is the edge order, as before).
  struct StaticSeg {
    std::string err;
    int pdmin = 0, pdmax = 0, nt = 0, rs = 0;
    std::vector<uint32_t> seg, ew;
    std::vector<long long> eq;
  };
  std::vector<StaticSeg> sseg(d->n_projs);
  std::vector<long long> gid0_of(d->n_pops);
  for (int p = 0; p < d->n_pops; ++p) gid0_of[p] = net->pops[p].gid0;
  {
    auto build_one = [&](int q) {
      const bs_projection& pj = d->projs[q];
      StaticSeg& S = sseg[q];
      if (pj.plastic || pj.on_device || pj.pre_pop < 0 || pj.pre_pop >= d->n_pops || pj.post_pop < 0 ||
          pj.post_pop >= d->n_pops || pj.channel < 0 || pj.channel >= d->n_channels)
        return;  // plastic, device edges, or rejected with its message by the serial pass
      const int kind = d->channels[pj.channel].kind;
      if ((kind != BS_CH_COUNT && kind != BS_CH_FIXED) || (kind == BS_CH_FIXED && !pj.weight) ||
          pj.n_edges < 0 || pj.n_edges >= (1ll << 32))
        return;
      const int npre = d->pops[pj.pre_pop].n, npost = d->pops[pj.post_pop].n;
      const int nt = (npost + TS - 1) / TS;
      const long long g0 = gid0_of[pj.pre_pop];
      int pdmin = 1 << 30, pdmax = 0;
      std::vector<uint32_t> cnt((size_t)npre * (nt + 1) + 1, 0);
      for (long long e = 0; e < pj.n_edges; ++e) {
        const int a = pj.pre[e], b = pj.post[e];
        if (a < 0 || a >= npre || b < 0 || b >= npost) {
          S.err = "edge index out of range";
          return;
        }
        const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
        if (dd < 0 || dd >= (1 << 24)) {
          S.err = "delay must be in [0, 2^24) steps";
          return;
        }
        pdmin = std::min(pdmin, dd);
        pdmax = std::max(pdmax, dd);
        cnt[(size_t)a * (nt + 1) + b / TS + 1]++;
        __atomic_fetch_add(&net->outdeg[g0 + a], 1ll, __ATOMIC_RELAXED);
      }
      if (pj.n_edges == 0) pdmin = pdmax = 0;
      // edge words delay * RS + lane, RS the post population's arrival-counter step stride
      const int RS = (kind == BS_CH_COUNT ? net->pops[pj.post_pop].n_cw : net->pops[pj.post_pop].n_fx) * TS;
      if ((unsigned long long)pdmax * (unsigned long long)RS + (unsigned long long)TS > 0xFFFFFFFFull) {
        S.err = "delay too long for this population's arrival counters (delay * stride >= 2^32)";
        return;
      }
      S.rs = RS;
      // exclusive prefix over (pre, tile) in row-major order; entry (a, nt) is the row end
      S.seg.assign((size_t)npre * (nt + 1), 0);
      uint32_t acc = 0;
      for (int a = 0; a < npre; ++a) {
        for (int t = 0; t < nt; ++t) {
          S.seg[(size_t)a * (nt + 1) + t] = acc;
          acc += cnt[(size_t)a * (nt + 1) + t + 1];
        }
        S.seg[(size_t)a * (nt + 1) + nt] = acc;
      }
      std::vector<uint32_t> pos(S.seg);
      S.ew.assign(pj.n_edges + 4, 0u);  // the gather reads whole 16-byte chunks past a segment's end
      if (kind == BS_CH_FIXED) S.eq.assign(pj.n_edges, 0);
      const int F = d->channels[pj.channel].fixed_shift;
      for (long long e = 0; e < pj.n_edges; ++e) {
        const int a = pj.pre[e], b = pj.post[e];
        const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
        const uint32_t k = pos[(size_t)a * (nt + 1) + b / TS]++;
        S.ew[k] = (uint32_t)dd * (uint32_t)RS + (uint32_t)(b % TS);
        if (kind == BS_CH_FIXED) S.eq[k] = (long long)std::llrint(std::ldexp(pj.weight[e], F));
      }
      S.pdmin = pdmin;
      S.pdmax = pdmax;
      S.nt = nt;
    };
    const int nthr = (int)std::max(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(nthr, d->n_projs); ++t)
      pool.emplace_back([&] {
        for (int q = next++; q < d->n_projs; q = next++) build_one(q);
      });
    for (std::thread& th : pool) th.join();
  }
  SegScratch seg_scratch;  // device segment builds (projections with on_device edges)
  for (int q = 0; q < d->n_projs; ++q) {
    const bs_projection& pj = d->projs[q];
    if (pj.pre_pop < 0 || pj.pre_pop >= d->n_pops || pj.post_pop < 0 || pj.post_pop >= d->n_pops)
      BAD("projection population index out of range");
    if (pj.channel < 0 || pj.channel >= d->n_channels || chan_pop[pj.channel] != pj.post_pop)
      BAD("projection channel does not belong to the post population");
    if (chan_proj[pj.channel] >= 0) BAD("two projections share one channel");
    chan_proj[pj.channel] = q;
    net->proj_stdp.push_back(-1);
    const int kind = d->channels[pj.channel].kind;
    if (pj.plastic) {
      if (pj.on_device) BAD("a plastic projection takes host edge arrays (on_device = 0)");
      if (d->pops[pj.pre_pop].remote || d->pops[pj.post_pop].remote)
        BAD("a plastic projection needs both populations on this shard");
      if (kind != BS_CH_FIXED) BAD("a plastic projection needs a FIXED channel");
      if (!pj.weight) BAD("a plastic projection needs per-edge initial weights");
      if (!(pj.stdp.Wmax != 0.0) || !(pj.stdp.tau_plus > 0) || !(pj.stdp.tau_minus > 0))
        BAD("invalid stdp_synapse parameters");
      if (pj.n_edges < 0 || pj.n_edges >= (1ll << 32)) BAD("projection edge count out of range");
      DevPop& post_dp = net->pops[pj.post_pop];
      DevPop& pre_dp = net->pops[pj.pre_pop];
      if (post_dp.n_pl >= BS_MAX_PL || pre_dp.n_plpre >= BS_MAX_PL)
        BAD("more than 4 plastic projections into or out of one population");
      const int npre = d->pops[pj.pre_pop].n, npost = d->pops[pj.post_pop].n;
      const int nt = (npost + TS - 1) / TS;
      const long long E = pj.n_edges;
      int pdmin = 1 << 30, pdmax = 0;
      // stored order: stable by (pre, post tile), the gather segments; reference order: stable by pre
      std::vector<uint32_t> seg((size_t)npre * (nt + 1) + 1, 0);
      for (long long e = 0; e < E; ++e) {
        const int a = pj.pre[e], b = pj.post[e];
        if (a < 0 || a >= npre || b < 0 || b >= npost) BAD("edge index out of range");
        const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
        if (dd < 1) BAD("plastic projections need delays of at least one step");
        if (dd >= (1 << 24)) BAD("delay must be in [0, 2^24) steps");
        pdmin = std::min(pdmin, dd);
        pdmax = std::max(pdmax, dd);
        seg[(size_t)a * (nt + 1) + b / TS + 1]++;
        net->outdeg[net->pops[pj.pre_pop].gid0 + a]++;
      }
      if (E == 0) pdmin = pdmax = 1;
      {  // exclusive prefix over (pre, tile); entry (a, nt) is the row end
        uint32_t acc = 0;
        for (int a = 0; a < npre; ++a) {
          for (int t = 0; t < nt; ++t) {
            const uint32_t c = seg[(size_t)a * (nt + 1) + t + 1];
            seg[(size_t)a * (nt + 1) + t] = acc;
            acc += c;
          }
          seg[(size_t)a * (nt + 1) + nt] = acc;
        }
        seg.resize((size_t)npre * (nt + 1));
      }
      std::vector<long long> spos(E);
      std::vector<uint32_t> ew(E);
      std::vector<int> epre(E), epost(E);
      std::vector<double> w(E);
      {
        std::vector<uint32_t> cur(seg);
        for (long long e = 0; e < E; ++e) {
          const int a = pj.pre[e], b = pj.post[e];
          const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
          const uint32_t k = cur[(size_t)a * (nt + 1) + b / TS]++;
          spos[e] = k;
          ew[k] = ((uint32_t)dd << 8) | (uint32_t)(b % TS);
          epre[k] = a;
          epost[k] = b;
          w[k] = pj.weight[e];
        }
      }
      std::vector<long long> ref(E);
      for (long long e = 0; e < E; ++e) ref[e] = e;
      std::stable_sort(ref.begin(), ref.end(), [&](long long x, long long y) { return pj.pre[x] < pj.pre[y]; });
      std::vector<long long> wref(E);
      for (long long k = 0; k < E; ++k) wref[k] = spos[ref[k]];
      std::vector<unsigned> cptr(npost + 1, 0), ceid(E);
      for (long long k = 0; k < E; ++k) cptr[epost[k] + 1]++;
      for (int b = 0; b < npost; ++b) cptr[b + 1] += cptr[b];
      {
        std::vector<unsigned> pos(cptr.begin(), cptr.end() - 1);
        for (long long k = 0; k < E; ++k) ceid[pos[epost[k]]++] = (unsigned)k;
      }
      StdpDev S;
      memset(&S, 0, sizeof(S));
      S.pre_pop = pj.pre_pop;
      S.post_pop = pj.post_pop;
      S.pre_gid0 = net->pops[pj.pre_pop].gid0;
      S.post_gid0 = net->pops[pj.post_pop].gid0;
      S.n_pre = npre;
      S.n_post = npost;
      S.d_min = pdmin;
      S.d_max = pdmax;
      S.nt = nt;
      S.shift = d->channels[pj.channel].fixed_shift;
      S.decay_plus = pj.stdp.decay_plus;
      S.decay_minus = pj.stdp.decay_minus;
      S.lam = pj.stdp.lambda_;
      S.alpha = pj.stdp.alpha;
      S.mup = pj.stdp.mu_plus;
      S.mum = pj.stdp.mu_minus;
      S.Wmax = pj.stdp.Wmax;
      uint32_t *dseg, *dew;
      int *dpre, *dla;
      unsigned *dcptr, *dceid;
      double *dw, *dV;
      ew.resize(E + 1, 0u);  // no zero-length device arrays
      epre.resize(E + 1, 0);
      ceid.resize(E + 1, 0u);
      w.resize(E + 1, 0.0);
      CHK(upload(net, &dseg, seg.data(), seg.size()));
      CHK(upload(net, &dew, ew.data(), ew.size()));
      CHK(upload(net, &dpre, epre.data(), epre.size()));
      CHK(upload(net, &dcptr, cptr.data(), cptr.size()));
      CHK(upload(net, &dceid, ceid.data(), ceid.size()));
      CHK(upload(net, &dw, w.data(), w.size()));
      w.resize(E);
      CHK(dmalloc(&dla, (size_t)E + 1));
      net->owned.push_back(dla);
      CHK(dmalloc(&dV, (size_t)std::max(1, npost)));
      net->owned.push_back(dV);
      S.seg = dseg;
      S.ew = dew;
      S.e_pre = dpre;
      S.csc_ptr = dcptr;
      S.csc_eid = dceid;
      S.w = dw;
      S.la = dla;
      S.V = dV;  // U: allocated once the window length (and with it Lu) is known
      const int qs = (int)net->stdp.size();
      net->chans[pj.channel].plastic = post_dp.n_pl;
      post_dp.pl[post_dp.n_pl++] = qs;
      pre_dp.plpre[pre_dp.n_plpre++] = qs;
      net->proj_stdp.back() = qs;
      net->stdp.push_back(S);
      net->w0.push_back(w);
      net->w_ref.push_back(wref);
      dmin = std::min(dmin, pdmin);
      dmax = std::max(dmax, pdmax);
      continue;
    }
    if (kind != BS_CH_COUNT && kind != BS_CH_FIXED) BAD("projection channel must be COUNT or FIXED");
    if (kind == BS_CH_FIXED && !pj.weight) BAD("FIXED projection needs per-edge weights");
    if (pj.n_edges < 0 || pj.n_edges >= (1ll << 32)) BAD("projection edge count out of range");
    GProj G;
    memset(&G, 0, sizeof(G));
    G.pre_pop = pj.pre_pop;
    G.chan = pj.channel;
    int pdmin = 0, pdmax = 0;
    if (pj.on_device) {
      const int RS = (kind == BS_CH_COUNT ? net->pops[pj.post_pop].n_cw : net->pops[pj.post_pop].n_fx) * TS;
      CHK(build_segments_device(net, pj, kind, d->channels[pj.channel].fixed_shift, RS, d->pops[pj.pre_pop].n,
                                d->pops[pj.post_pop].n, gid0_of[pj.pre_pop], seg_scratch, G, pdmin, pdmax));
      G.rs = RS;
      G.ne = pj.n_edges;
    } else {
      StaticSeg& S = sseg[q];
      G.ne = pj.n_edges;
      if (!S.err.empty()) BAD(S.err);
      pdmin = S.pdmin;
      pdmax = S.pdmax;
      G.nt = S.nt;
      G.rs = S.rs;
      uint32_t *dseg = nullptr, *dew = nullptr;
      long long* deq = nullptr;
      CHK(upload(net, &dseg, S.seg.data(), S.seg.size()));
      CHK(upload(net, &dew, S.ew.data(), S.ew.size()));
      if (kind == BS_CH_FIXED) CHK(upload(net, &deq, S.eq.data(), S.eq.size()));
      std::vector<uint32_t>().swap(S.seg);  // host copies no longer needed
      std::vector<uint32_t>().swap(S.ew);
      std::vector<long long>().swap(S.eq);
      G.seg = dseg;
      G.edges = dew;
      G.q = deq;
    }
    G.d_min = pdmin;
    G.d_max = pdmax;
    pop_gp[pj.post_pop].push_back((int)gp_h.size());
    gp_h.push_back(G);
    if (pj.n_edges) {
      dmin = std::min(dmin, pdmin);
      dmax = std::max(dmax, pdmax);
    }
  }
  if (dmin > dmax) dmin = dmax = 0;
  for (int c = 0; c < d->n_channels; ++c) {
    const int k = d->channels[c].kind;
    if ((k == BS_CH_COUNT || k == BS_CH_FIXED) && chan_proj[c] < 0)
      BAD("COUNT/FIXED channel without a projection");
  }
  // gather list in post-population order
  std::vector<GProj> gp_sorted;
  for (int p = 0; p < d->n_pops; ++p) {
    net->pops[p].gp_begin = (int)gp_sorted.size();
    net->pops[p].gp_count = (int)pop_gp[p].size();
    for (int k : pop_gp[p]) gp_sorted.push_back(gp_h[k]);
  }
  net->d_min = dmin;
  net->d_max = dmax;
  // 16-bit edge words when every static projection's words (delay * stride +
  // lane) fit and the (pre, tile) segments are long enough for the halved
  // bytes and chunk rounds to pay for the unpacking: measured on one box,
  // +4 % on the 256-network ensemble (12.8 edges per segment) and -3 % on
  // Brunel-11250 (3.2 per segment with 32-neuron tiles). The 32-bit arrays
  // are converted on the device and released.
  {
    bool nar = !gp_sorted.empty();
    double ne = 0.0, nseg = 0.0;
    for (const GProj& G : gp_sorted) {
      if ((long long)G.d_max * G.rs + TS > 65535) nar = false;
      ne += (double)G.ne;
      nseg += (double)net->pops[G.pre_pop].n * G.nt;
    }
    if (ne < 6.0 * nseg) nar = false;
    if (nar) {
      for (GProj& G : gp_sorted) {
        uint16_t* dn = nullptr;
        CHK(dmalloc(&dn, (size_t)G.ne + 8));
        net->owned.push_back(dn);
        hipLaunchKernelGGL(k_narrow_edges, dim3((unsigned)std::min<long long>((G.ne + 8 + 255) / 256, 65535)),
                           dim3(256), 0, net->stream, G.edges, G.ne, dn);
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(net->stream));
        void* old = (void*)G.edges;
        net->owned.erase(std::remove(net->owned.begin(), net->owned.end(), old), net->owned.end());
        HCHK(hipFree(old));
        G.edges = (const uint32_t*)dn;
      }
      net->narrow = true;
    }
  }

  // ---- Poisson producers: when a launch of one or two rounds of tiles leaves
  // at least 8 block slots free, blocks in those slots draw each population's
  // first Poisson channel one window ahead (counts up to 65535) and the tiles
  // read the counts instead
  net->pop_pc.assign(net->pops.size(), 0);
  {
    int n_cu = 0;
    HCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, net->device));
    // blocks beyond the tiles on the CUs of a launch's one or two rounds
    const long long nt_ = (long long)net->tiles.size();
    const long long slots = nt_ <= n_cu ? n_cu : (nt_ <= 2ll * n_cu ? 2ll * n_cu : nt_);
    const long long spare = slots - nt_;
    if (spare >= 8) {
      for (size_t p = 0; p < net->pops.size(); ++p) {
        const DevPop& dp = net->pops[p];
        for (int c = dp.ch_begin; c < dp.ch_begin + dp.ch_count; ++c) {
          DevChan& dc = net->chans[c];
          if (dc.kind != BS_CH_POISSON || dc.frac != 0.0) continue;  // lerped channels draw in place
          if ((long long)(dc.cdf_len - 1) * dc.parts > 65535) continue;
          dc.pre = 1;
          net->pop_pc[p] = 1;
          net->pjobs.push_back(PJob{dp.gid0, dp.n, c});
          net->pjob_n += dp.n;
          break;
        }
      }
      if (!net->pjobs.empty()) net->n_prod = (int)std::min<long long>(spare, 256);
    }
  }

  // ---- window length: W = d_min + 1 (holder lag) capped by the LDS budget
  int W = dmin + 1;
  const int maxw = d->max_window > 0 ? d->max_window : 32;
  if (W > maxw) W = maxw;
  if (W < 1) W = 1;
  // tail channels: static delta channels summed after a plastic channel of
  // the same receptor wait in LDS for the update loop
  for (size_t q = 0; q < net->pops.size(); ++q) {
    DevPop& p = net->pops[q];
    for (int rc = 0; rc < 2; ++rc) {
      bool seen = false;
      for (int c = p.ch_begin; c < p.ch_begin + p.ch_count; ++c) {
        DevChan& dc = net->chans[c];
        if (dc.receptor != rc) continue;
        if (dc.plastic >= 0) seen = true;
        else if (seen && (dc.kind == BS_CH_COUNT || dc.kind == BS_CH_FIXED || dc.kind == BS_CH_POISSON))
          dc.tail = p.n_tail++;
      }
    }
  }
  // LDS budget: 64 KB (two blocks per CU), or with plastic projections most
  // of a CU's LDS, for the window's arrival records
  size_t lds_max = kLdsBudget;  // the kernels' dynamic-LDS attribute (one value for every network)
  {
    int mx = 0;
    HCHK(hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, net->device));
    lds_max = std::max<size_t>(kLdsBudget, std::min<size_t>(mx > 4096 ? (size_t)mx - 4096 : 0, 150 * 1024));
  }
  const size_t budget = net->stdp.empty() ? kLdsBudget : lds_max;
  net->lds_budget = lds_max;
  int pl_span = 0;  // source steps of a plastic gather, beyond the window's W
  for (const StdpDev& S : net->stdp) pl_span = std::max(pl_span, S.d_max - S.d_min);
  const int kPlMin = 512;  // arrival records a sub-window must be able to hold
  auto pl_end = [&](const DevPop& p, int w, size_t base, int cap) {
    if (p.n_pl == 0 && p.n_plpre == 0) return base;
    return pl_layout(base, w, TS, p.n_pl, pl_span + w, p.n_tail, (int)net->tsize, p.n_pl ? cap : 0).end;
  };
  // LDS layout of k_window (win_layout), first without staged tables
  // each window gathers the spikes of the previous one (at most W steps)
  auto span_of = [&](int w) { return w; };
  // LDS counters for every arrival offset 0..d_max when they fit (the window's
  // later arrivals are then added to the ring with plain stores), else only
  // for the window's own W steps (later arrivals go to the ring atomically)
  bool full_h = true;
  auto hz = [&](int w) { return full_h ? std::max(w, dmax + 1) : w; };
  auto lds_of = [&](int w) {
    size_t best = 0;
    const int SP = span_of(w) + 1;
    for (size_t q = 0; q < net->pops.size(); ++q) {
      const DevPop& p = net->pops[q];
      best = std::max(best, pl_end(p, w, win_layout(w, hz(w), TS, p.n_cw, p.n_fx, p.gp_count, SP, p.ch_count,
                                       (p.model == BS_AEIF_COND_ALPHA ? 2 : 1) + (p.n_cur > 0 ? 1 : 0), (int)net->tsize, 0,
                                       net->pop_pc[q]).end, kPlMin));
    }
    return best;
  };
  if (lds_of(W) > budget) full_h = false;
  while (W > 1 && lds_of(W) > budget) --W;
  if (lds_of(W) > budget) BAD("too many input channels per population for the LDS budget");
  net->W = W;
  net->H = hz(W);
  net->max_span = span_of(W);
  net->psp = pl_span + W;
  // stage count tables and Poisson CDFs into LDS while they fit (leaving a
  // plastic population room for its minimum of arrival records)
  {
    const int SP = net->max_span + 1;
    size_t need = 0;
    for (size_t q = 0; q < net->pops.size(); ++q) {
      DevPop& p = net->pops[q];
      const size_t base_end = pl_end(p, W, win_layout(W, net->H, TS, p.n_cw, p.n_fx, p.gp_count, SP, p.ch_count,
                                         (p.model == BS_AEIF_COND_ALPHA ? 2 : 1) + (p.n_cur > 0 ? 1 : 0), (int)net->tsize, 0,
                                         net->pop_pc[q]).end, kPlMin);
      size_t off = 0;
      for (int c = p.ch_begin; c < p.ch_begin + p.ch_count; ++c) {
        DevChan& dc = net->chans[c];
        dc.lds_off = -1;
        size_t sz = 0;
        // a count table is staged up to S[255]: per step and neuron larger
        // counts are rare, and those read the table in global memory
        if (dc.kind == BS_CH_COUNT) {
          dc.lds_len = std::min(dc.max_count + 1, kStageCount);
          sz = al8((size_t)dc.lds_len * net->tsize);
        }
        else if (dc.kind == BS_CH_POISSON) sz = al8((size_t)dc.cdf_len * 8 + 64 * 2);
        if (sz && base_end + off + sz + 16 <= budget) {
          dc.lds_off = (long long)off;
          off += sz;
        }
      }
      p.tab_bytes = (long long)off;
      const size_t wend = win_layout(W, net->H, TS, p.n_cw, p.n_fx, p.gp_count, SP, p.ch_count,
                                     (p.model == BS_AEIF_COND_ALPHA ? 2 : 1) + (p.n_cur > 0 ? 1 : 0), (int)net->tsize,
                                     (size_t)off, net->pop_pc[q]).end;
      if (p.n_pl > 0) {  // the rest of the budget holds arrival records
        const size_t e0 = pl_end(p, W, wend, 0);
        const long long room = ((long long)budget - (long long)e0 - 64) / 12;
        p.pl_cap = (int)std::max<long long>(kPlMin, std::min<long long>(room, 16384));
        if (d->plastic_records > 0) p.pl_cap = std::min(p.pl_cap, (int)d->plastic_records);
      }
      need = std::max(need, pl_end(p, W, wend, p.pl_cap));
    }
    if (need > budget) BAD("LDS layout exceeds the budget");
    net->lds_bytes = need;
  }
  for (StdpDev& S : net->stdp) {  // pre-trace rings: U_i(m) of m in [n0 - d_max, n0 + W]
    S.Lu = S.d_max + W + 2;
    double* dU = nullptr;
    CHK(dmalloc(&dU, (size_t)S.Lu * std::max(1, S.n_pre)));
    net->owned.push_back(dU);
    S.U = dU;
  }
  net->Ls = dmax + 2 * W + 2;

  // ---- per-population blobs (DevPop, channels, gather sources, staged
  // tables in the LDS layout) and the complete tile records
  {
    std::vector<unsigned char>& B = net->h_blob;
    std::vector<long long> boff(net->pops.size());
    for (size_t p = 0; p < net->pops.size(); ++p) {
      const DevPop& dp = net->pops[p];
      const BlobLayout L = blob_layout(dp.ch_count, dp.gp_count, (size_t)dp.tab_bytes);
      const size_t o = al16(B.size());
      B.resize(o + L.end, 0);
      boff[p] = (long long)o;
      memcpy(&B[o + L.pop], &dp, sizeof(DevPop));
      if (dp.ch_count)
        memcpy(&B[o + L.ch], &net->chans[dp.ch_begin], sizeof(DevChan) * dp.ch_count);
      for (int g = 0; g < dp.gp_count; ++g) {
        const GProj& G = gp_sorted[dp.gp_begin + g];
        const DevChan& ch = net->chans[G.chan];
        GInfo I;
        memset(&I, 0, sizeof(I));
        I.seg = G.seg;
        I.edges = G.edges;
        I.q = G.q;
        I.pre_gid0 = net->pops[G.pre_pop].gid0;
        I.nt1 = G.nt + 1;
        I.kind = ch.kind;
        I.cw = ch.cw;
        I.half = ch.half;
        I.fx = ch.fx;
        I.rs = G.rs;
        memcpy(&B[o + L.gi + g * sizeof(GInfo)], &I, sizeof(GInfo));
      }
      for (int c = dp.ch_begin; c < dp.ch_begin + dp.ch_count; ++c) {
        const DevChan& ch = net->chans[c];
        if (ch.lds_off < 0) continue;
        unsigned char* dst = &B[o + L.tab + (size_t)ch.lds_off];
        if (ch.kind == BS_CH_COUNT) {
          for (int k = 0; k < ch.lds_len; ++k) {
            const double v = d->tables[ch.table_off + k];
            if (net->tsize == 8) memcpy(dst + 8 * k, &v, 8);
            else {
              const float f = (float)v;
              memcpy(dst + 4 * k, &f, 4);
            }
          }
        } else if (ch.kind == BS_CH_POISSON) {
          memcpy(dst, net->h_cdf[c].data(), 8 * net->h_cdf[c].size());
          memcpy(dst + 8 * net->h_cdf[c].size(), net->h_guide[c].data(), 2 * net->h_guide[c].size());
        }
      }
    }
    for (size_t t = 0; t < net->tiles.size(); ++t) {
      Tile& T = net->tiles[t];
      const DevPop& dp = net->pops[T.pop];
      T.tile_in_pop = (int)t - dp.tile0;
      T.ncw = dp.n_cw;
      T.nfx = dp.n_fx;
      T.ngp = dp.gp_count;
      T.nch = dp.ch_count;
      T.nrec = dp.model == BS_AEIF_COND_ALPHA ? 2 : 1;
      T.blob_words = (int)(blob_layout(dp.ch_count, dp.gp_count, (size_t)dp.tab_bytes).end / 16);
      T.gp_begin = dp.gp_begin;
      T.pois = net->pop_pc[T.pop];
      T.ncur = dp.n_cur > 0 ? 1 : 0;
      for (int g = 0; g < 4; ++g) T.pre_pop[g] = g < dp.gp_count ? gp_sorted[dp.gp_begin + g].pre_pop : 0;
      T.gid0 = dp.gid0;
      T.blob_off = boff[T.pop];
      T.tab_bytes = dp.tab_bytes;
    }
    net->pop_blob = boff;
  }

  // ---- launch order of the tiles. A launch with more tiles than two per CU
  // runs in rounds. Block b runs on XCD b % 8 and k_window gives XCD x the
  // x-th contiguous share of the tile array, so the shares are filled here
  // with runs of up to 16 tiles of one population, heaviest first, each run
  // on the XCD with the least estimated work so far: the XCDs finish
  // together and their last blocks are light ones. Estimated work of a tile:
  // its neurons times (64 + the population's in-degree). The order changes no
  // result (every sum is order independent).
  {
    int n_cu = 0;
    HCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, net->device));
    const int G = (int)net->tiles.size();
    if (G > 2 * n_cu) {
      std::vector<double> indeg(net->pops.size(), 0.0);
      for (int q = 0; q < d->n_projs; ++q) {
        const bs_projection& pj = d->projs[q];
        const int npost = net->pops[pj.post_pop].n;
        if (npost > 0) indeg[pj.post_pop] += (double)pj.n_edges / npost;
      }
      struct Run {
        int first, len;
        double w;  // per tile
      };
      std::vector<Run> runs;
      for (int t = 0; t < G;) {
        const int p = net->tiles[t].pop;
        int e = t;
        while (e < G && e - t < 16 && net->tiles[e].pop == p) ++e;
        runs.push_back(Run{t, e - t, (double)net->tiles[t].count * (64.0 + indeg[p])});
        t = e;
      }
      std::stable_sort(runs.begin(), runs.end(), [](const Run& x, const Run& y) { return x.w > y.w; });
      const int q8 = G / 8, rem = G % 8;
      int cap[8];
      double load[8];
      std::vector<int> lists[8];
      for (int x = 0; x < 8; ++x) {
        cap[x] = q8 + (x < rem ? 1 : 0);
        load[x] = 0.0;
      }
      for (const Run& r : runs) {
        int first = r.first, left = r.len;
        while (left > 0) {
          int best = -1;
          for (int x = 0; x < 8; ++x)
            if (cap[x] > 0 && (best < 0 || load[x] < load[best])) best = x;
          const int take = std::min(left, cap[best]);
          for (int k = 0; k < take; ++k) {
            const Tile& Tk = net->tiles[first + k];
            lists[best].push_back(first + k);
            load[best] += (double)Tk.count * (64.0 + indeg[Tk.pop]);
          }
          cap[best] -= take;
          first += take;
          left -= take;
        }
      }
      std::vector<Tile> order;
      order.reserve(G);
      for (int x = 0; x < 8; ++x)
        for (int t : lists[x]) order.push_back(net->tiles[t]);
      net->tiles.swap(order);
    }
  }

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
  CHK(dmalloc(&net->d_chans, net->chans.size()));
  CHK(dmalloc(&net->d_gproj, gp_sorted.size()));
  CHK(dmalloc((char**)&net->d_tables, std::max<long long>(1, d->n_tables) * ts));
  CHK(dmalloc((char**)&net->d_state, (size_t)BS_NSTATE * gid * ts));
  CHK(dmalloc(&net->d_ref, gid));
  CHK(dmalloc(&net->d_sb, (size_t)net->Ls * gid));
  CHK(dmalloc(&net->d_sb_count, (size_t)net->Ls * d->n_pops));
  CHK(dmalloc(&net->d_rec_count, 2));
  CHK(dmalloc(&net->d_flags, 1));
  CHK(dmalloc(&net->d_base, 1));
  CHK(dmalloc(&net->d_prevlen, 1));
  CHK(dmalloc(&net->d_blob, std::max<size_t>(16, net->h_blob.size())));
  HCHK(hipMemcpy(net->d_blob, net->h_blob.data(), net->h_blob.size(), hipMemcpyHostToDevice));
  {  // arrival ring: slots for steps n0 .. n0 + d_max
    net->Lr = net->d_max + 1;
    int ncw = 1, nfx = 0;
    for (const DevPop& p : net->pops) {
      ncw = std::max(ncw, p.n_cw);
      nfx = std::max(nfx, p.n_fx);
    }
    net->NCW = ncw;
    net->NFX = nfx;
    CHK(dmalloc(&net->d_ring_cnt, (size_t)net->Lr * ncw * std::max<long long>(1, gid)));
    CHK(dmalloc(&net->d_ring_fx, nfx > 0 ? (size_t)net->Lr * nfx * std::max<long long>(1, gid) : 1));
  }
  if (!net->stdp.empty()) {
    CHK(dmalloc(&net->d_spk_hist, (size_t)net->Ls * gid));
    CHK(dmalloc(&net->d_stdp, net->stdp.size()));
    HCHK(hipMemcpy(net->d_stdp, net->stdp.data(), sizeof(StdpDev) * net->stdp.size(),
                   hipMemcpyHostToDevice));
  }
  CHK(dmalloc(&net->d_vm_col, gid));
  if (net->n_prod > 0) {
    CHK(dmalloc(&net->d_pcnt, (size_t)2 * W * std::max<long long>(1, gid)));
    CHK(dmalloc(&net->d_pjobs, net->pjobs.size()));
    HCHK(hipMemcpy(net->d_pjobs, net->pjobs.data(), sizeof(PJob) * net->pjobs.size(), hipMemcpyHostToDevice));
    // about four draws per producer thread and window
    const long long want = (W * net->pjob_n + 1023) / 1024;
    net->n_prod = (int)std::max<long long>(1, std::min<long long>(net->n_prod, want));
  }
  {  // spike record list: worst case every neuron fires every step of K windows
    const long long per_win = std::max<long long>(1, gid) * W;
    long long k = (256ll << 20) / per_win;
    if (k < 1) k = 1;
    if (k > 512) k = 512;
    net->rec_cap = per_win * k;
    CHK(dmalloc(&net->d_rec[0], net->rec_cap));
    CHK(dmalloc(&net->d_rec[1], net->rec_cap));
    CHK(dmalloc(&net->d_rec_sorted, net->rec_cap));
    net->vm_cap_steps = k * W;
    while ((1ll << net->gid_bits) < gid) ++net->gid_bits;
    while ((1ll << net->step_bits) < net->vm_cap_steps) ++net->step_bits;
    HCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, net->sort_tmp_bytes, net->d_rec[0], net->d_rec_sorted,
                                           (int)net->rec_cap, 0, net->gid_bits + net->step_bits));
    {
      void* t = nullptr;
      HCHK(hipMalloc(&t, std::max<size_t>(net->sort_tmp_bytes, 256)));
      net->d_sort_tmp = t;
    }
    CHK(dmalloc(&net->d_vm, (size_t)std::max(1, ncol) * net->vm_cap_steps));
  }
  HCHK(hipMemcpy(net->d_tiles, net->tiles.data(), sizeof(Tile) * net->tiles.size(), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(net->d_pops, net->pops.data(), sizeof(DevPop) * net->pops.size(), hipMemcpyHostToDevice));
  if (!net->chans.empty())
    HCHK(hipMemcpy(net->d_chans, net->chans.data(), sizeof(DevChan) * net->chans.size(), hipMemcpyHostToDevice));
  if (!gp_sorted.empty())
    HCHK(hipMemcpy(net->d_gproj, gp_sorted.data(), sizeof(GProj) * gp_sorted.size(), hipMemcpyHostToDevice));
  if (d->n_tables > 0) {
    if (ts == 8) {
      HCHK(hipMemcpy(net->d_tables, d->tables, 8 * d->n_tables, hipMemcpyHostToDevice));
    } else {
      std::vector<float> tf(d->n_tables);
      for (long long i = 0; i < d->n_tables; ++i) tf[i] = (float)d->tables[i];
      HCHK(hipMemcpy(net->d_tables, tf.data(), 4 * d->n_tables, hipMemcpyHostToDevice));
    }
  }
  if (gid) HCHK(hipMemcpy(net->d_vm_col, vm_col.data(), 4 * gid, hipMemcpyHostToDevice));
  CHK(dmalloc(&net->d_tap_col, std::max<size_t>(1, net->vm_tap_col.size())));
  if (!net->vm_tap_col.empty())
    HCHK(hipMemcpy(net->d_tap_col, net->vm_tap_col.data(), 4 * net->vm_tap_col.size(), hipMemcpyHostToDevice));
  {
    const void* kw[8] = {(const void*)k_window<double, false, false>, (const void*)k_window<float, false, false>,
                         (const void*)k_window<double, true, false>,  (const void*)k_window<float, true, false>,
                         (const void*)k_window<double, false, true>,  (const void*)k_window<float, false, true>,
                         (const void*)k_window<double, true, true>,   (const void*)k_window<float, true, true>};
    for (const void* f : kw)
      HCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)net->lds_budget));
  }
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
  HIPCHK(hipMemset(net->d_sb_count, 0, 4 * (size_t)net->Ls * net->pops.size()));
  HIPCHK(hipMemset(net->d_flags, 0, sizeof(int)));
  for (size_t q = 0; q < net->stdp.size(); ++q) {  // traces 0, no arrivals yet, initial weights
    const StdpDev& S = net->stdp[q];
    HIPCHK(hipMemset(S.U, 0, 8 * (size_t)S.Lu * std::max(1, S.n_pre)));
    HIPCHK(hipMemset(S.V, 0, 8 * (size_t)std::max(1, S.n_post)));
    HIPCHK(hipMemset(S.la, 0xFF, 4 * (net->w0[q].size() + 1)));
    if (!net->w0[q].empty())
      HIPCHK(hipMemcpy(S.w, net->w0[q].data(), 8 * net->w0[q].size(), hipMemcpyHostToDevice));
  }
  if (net->d_spk_hist) HIPCHK(hipMemset(net->d_spk_hist, 0, (size_t)net->Ls * gid));
  HIPCHK(hipMemset(net->d_ring_cnt, 0, 4 * (size_t)net->Lr * net->NCW * std::max<long long>(1, gid)));
  if (net->NFX > 0)
    HIPCHK(hipMemset(net->d_ring_fx, 0, 8 * (size_t)net->Lr * net->NFX * std::max<long long>(1, gid)));
  HIPCHK(hipDeviceSynchronize());
  net->step = 0;
  net->last_win = 0;
  return BS_OK;
}

int bs_net_reset(bs_net* net) {
  if (!net) return fail(BS_EINVAL, "bs_net_reset: null handle");
  net->n_out_blocks = 0;
  net->n_out = 0;
  return upload_state(net);
}

int bs_net_step(bs_net* net, int64_t* step) {
  if (!net || !step) return fail(BS_EINVAL, "bs_net_step: null argument");
  *step = net->step;
  return BS_OK;
}

template <typename T>
static void launch_window(bs_net* net, int n0, int W, long long vm_step0, int prevlen,
                          const long long* base_ptr = nullptr, int widx = 0) {
  WinArgs a;
  a.ring_cnt = net->d_ring_cnt;
  a.ring_fx = net->d_ring_fx;
  a.Lr = net->Lr;
  a.H = net->H;
  a.NCW = net->NCW;
  a.NFX = net->NFX;
  a.prevlen = prevlen;
  a.prevlen_ptr = net->d_prevlen;
  a.base_ptr = base_ptr;
  a.widx = widx;
  a.spk_hist = net->d_spk_hist;
  a.stdp = net->d_stdp;
  a.psp = net->psp;
  a.tiles = net->d_tiles;
  a.blob = net->d_blob;
  a.pops = net->d_pops;
  a.chans = net->d_chans;
  a.tables = net->d_tables;
  a.gproj = net->d_gproj;
  a.state = net->d_state;
  a.ref = net->d_ref;
  a.Ntot = net->Ntot;
  a.sb = net->d_sb;
  a.sb_count = net->d_sb_count;
  a.Ls = net->Ls;
  a.n_pops = (int)net->pops.size();
  a.n0 = n0;
  a.W = W;
  a.Wmax = net->W;
  a.max_span = net->max_span;
  a.TS = net->TS;
  a.rec = net->d_rec[net->rec_par];
  a.gid_bits = net->gid_bits;
  a.rec_count = net->d_rec_count + net->rec_par;
  a.rec_cap = net->rec_cap;
  a.flags = net->d_flags;
  a.vm_col = net->d_vm_col;
  a.n_vm = net->n_vm_dev;
  a.vm_buf = net->d_vm;
  a.vm_step0 = vm_step0;
  a.n_tiles = (int)net->tiles.size();
  a.pcnt = net->d_pcnt;
  a.pjobs = net->d_pjobs;
  a.n_pjobs = (int)net->pjobs.size();
  a.pjob_n = net->pjob_n;
  const dim3 grid((unsigned)(net->tiles.size() + net->n_prod));
  const bool pl = !net->stdp.empty();
  if (net->narrow) {
    if (pl)
      hipLaunchKernelGGL((k_window<T, true, true>), grid, dim3(256), net->lds_bytes, net->stream, a);
    else
      hipLaunchKernelGGL((k_window<T, false, true>), grid, dim3(256), net->lds_bytes, net->stream, a);
  } else {
    if (pl)
      hipLaunchKernelGGL((k_window<T, true, false>), grid, dim3(256), net->lds_bytes, net->stream, a);
    else
      hipLaunchKernelGGL((k_window<T, false, false>), grid, dim3(256), net->lds_bytes, net->stream, a);
  }
}

// Poisson counts of the window starting at n0 (before a run's first window).
static void launch_pois_fill(bs_net* net, int n0) {
  WinArgs a;
  memset(&a, 0, sizeof(a));
  a.chans = net->d_chans;
  a.Ntot = net->Ntot;
  a.n0 = n0;
  a.Wmax = net->W;
  a.pcnt = net->d_pcnt;
  a.pjobs = net->d_pjobs;
  a.n_pjobs = (int)net->pjobs.size();
  a.pjob_n = net->pjob_n;
  hipLaunchKernelGGL(k_pois_fill, dim3((unsigned)net->n_prod), dim3(256), 0, net->stream, a);
}

// One step's weights of a plastic projection in the reference's CSR order.
__global__ void k_weight_tap(const double* __restrict__ w, const long long* __restrict__ ref, long long E,
                             double* __restrict__ row) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < E; k += (long long)gridDim.x * blockDim.x)
    row[k] = w[ref[k]];
}

int bs_net_set_timing(bs_net* net, int on) {
  if (!net) return fail(BS_EINVAL, "bs_net_set_timing: null handle");
  net->timing = on != 0;
  return BS_OK;
}

// One chunk of windows queued on the stream and not yet drained.
struct ChunkJob {
  long long c0 = 0, chunk = 0, done = 0;
  int par = 0, nwin = 0;
};

// The captured graph of one full chunk (nwin window launches, plus the STDP
// kernels of plastic networks) writing record buffer par; its kernels read the
// chunk's first step and the previous window's length from device memory.
static int chunk_graph(bs_net* net, int par, int nwin) {
  hipGraphExec_t& G = net->graph[par];
  if (G) return BS_OK;
  net->rec_par = par;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
  for (int w = 0; w < nwin; ++w) {
    if (net->tsize == 8)
      launch_window<double>(net, 0, net->W, 0, net->W, net->d_base, w);
    else
      launch_window<float>(net, 0, net->W, 0, net->W, net->d_base, w);
  }
  HIPCHK(hipStreamEndCapture(net->stream, &g));
  hipError_t e = hipGraphInstantiate(&G, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(BS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return BS_OK;
}

// Queues one chunk of windows writing record buffer j.par (a graph replay when
// the chunk is full), then the copies of its record count, the flags and, with
// pin_vm, its V_m rows into pinned memory, and marks its end with
// ev_chunk[j.par]. Returns after queueing; nothing here waits for the device.
static int enqueue_chunk(bs_net* net, const ChunkJob& j, bool pin_vm, std::vector<hipEvent_t>& evs,
                         double win_bytes, double step_bytes) {
  net->rec_par = j.par;
  HIPCHK(hipMemsetAsync(net->d_rec_count + j.par, 0, sizeof(unsigned long long), net->stream));
  if (net->timing) {
    for (size_t i = evs.size(); i < (size_t)j.nwin * 2; ++i) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      evs.push_back(e);
    }
  }
  const bool full = j.chunk == net->vm_cap_steps;
  if (full && !net->timing && net->wtaps.empty()) {
    // replay the captured chunk: K window launches reading the base step
    int rc = chunk_graph(net, j.par, j.nwin);
    if (rc != BS_OK) return rc;
    hipGraphExec_t& G = net->graph[j.par];
    // pinned slots of this buffer: rewritten two chunks later, after this
    // chunk's event has been waited on
    net->h_base_p[j.par] = j.c0;
    net->h_prev_p[j.par] = net->last_win;
    HIPCHK(hipMemcpyAsync(net->d_base, net->h_base_p + j.par, sizeof(long long), hipMemcpyHostToDevice,
                          net->stream));
    HIPCHK(hipMemcpyAsync(net->d_prevlen, net->h_prev_p + j.par, sizeof(int), hipMemcpyHostToDevice,
                          net->stream));
    HIPCHK(hipGraphLaunch(G, net->stream));
    net->last_win = net->W;
    net->n_update += j.nwin;
    net->b_update += win_bytes * j.nwin + step_bytes * j.chunk;
  } else {
    const int Wr = net->wtaps.empty() ? net->W : 1;  // record_weight taps: one step per launch
    for (int w = 0; w < j.nwin; ++w) {
      const long long n0 = j.c0 + (long long)w * Wr;
      const int Wc = (int)std::min<long long>(Wr, j.c0 + j.chunk - n0);
      if (net->timing) HIPCHK(hipEventRecord(evs[2 * w], net->stream));
      if (net->tsize == 8)
        launch_window<double>(net, (int)n0, Wc, j.c0, net->last_win);
      else
        launch_window<float>(net, (int)n0, Wc, j.c0, net->last_win);
      net->last_win = Wc;
      for (const bs_net::WTap& t : net->wtaps)  // the weights after this step
        if (t.E > 0)
          hipLaunchKernelGGL(k_weight_tap, dim3((unsigned)std::min<long long>(1024, (t.E + 255) / 256)), dim3(256), 0,
                             net->stream, (const double*)net->stdp[t.q].w, (const long long*)t.d_ref, t.E,
                             t.d_traj + (size_t)w * t.E);
      if (net->timing) HIPCHK(hipEventRecord(evs[2 * w + 1], net->stream));
      HIPCHK(hipGetLastError());
      net->n_update++;
      net->b_update += win_bytes + step_bytes * Wc;
    }
  }
  HIPCHK(hipMemcpyAsync(net->h_cnt + j.par, net->d_rec_count + j.par, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, net->stream));
  HIPCHK(hipMemcpyAsync(net->h_flags + j.par, net->d_flags, sizeof(int), hipMemcpyDeviceToHost, net->stream));
  if (pin_vm)
    HIPCHK(hipMemcpyAsync(net->h_vm_p + (size_t)j.par * net->vm_cap_steps * net->n_vm_dev, net->d_vm,
                          sizeof(double) * (size_t)j.chunk * net->n_vm_dev, hipMemcpyDeviceToHost, net->stream));
  HIPCHK(hipEventRecord(net->ev_chunk[j.par], net->stream));
  return BS_OK;
}

// Waits for chunk j, sorts its records by (step, gid) on the second stream,
// copies them into pinned memory and decodes them into the next spike block;
// fills the chunk's rows of vm_out. Runs while the next chunk computes.
static int drain_chunk(bs_net* net, const ChunkJob& j, bool pin_vm, double* vm_out,
                       std::vector<hipEvent_t>& evs) {
  HIPCHK(hipEventSynchronize(net->ev_chunk[j.par]));
  const int flags = net->h_flags[j.par];
  if (flags & FLAG_OVERFLOW) return fail(BS_ECAPACITY, "spike list overflow");
  if (flags & FLAG_COUNT) return fail(BS_ECAPACITY, "arrival event count exceeded the channel's in-degree");
  if (flags & FLAG_UNSTABLE) return fail(BS_ENUMERIC, "Numerical instability in aeif_cond_alpha dynamics.");
  if (flags & FLAG_PLASTIC)
    return fail(BS_ECAPACITY, "plastic arrivals of one step exceeded the window kernel's LDS record capacity");
  if (net->timing) {
    for (int w = 0; w < j.nwin; ++w) {
      float a = 0;
      HIPCHK(hipEventElapsedTime(&a, evs[2 * w], evs[2 * w + 1]));
      net->t_update += a;
    }
  }
  const unsigned long long cnt = net->h_cnt[j.par];
  if (cnt) {
    if (cnt > net->h_rec_cap) {
      // twice the count: a chunk a little busier than the largest so far
      // does not grow the buffer again (each growth pins new host memory)
      const size_t want = std::min<size_t>(std::max<size_t>(2 * (size_t)cnt, 2 * net->h_rec_cap),
                                           (size_t)std::max<long long>(net->rec_cap, (long long)cnt));
      if (net->h_rec) net->h_rec_old.push_back(net->h_rec);  // freed when the run ends
      net->h_rec = nullptr;
      net->h_rec_cap = 0;
      HIPCHK(hipHostMalloc((void**)&net->h_rec, sizeof(unsigned long long) * want, hipHostMallocDefault));
      net->h_rec_cap = want;
    }
    size_t tb = net->sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(net->d_sort_tmp, tb, net->d_rec[j.par], net->d_rec_sorted,
                                             (int)cnt, 0, net->gid_bits + net->step_bits, net->stream2));
    HIPCHK(hipMemcpyAsync(net->h_rec, net->d_rec_sorted, sizeof(unsigned long long) * cnt,
                          hipMemcpyDeviceToHost, net->stream2));
    HIPCHK(hipStreamSynchronize(net->stream2));
  }
  const unsigned long long gmask = (1ull << net->gid_bits) - 1;
  if (net->n_out_blocks == net->out_blocks.size()) net->out_blocks.emplace_back();
  bs_net::SpikeBlock& blk = net->out_blocks[net->n_out_blocks++];
  if (blk.gid.size() < cnt) {
    blk.gid.resize(cnt);
    blk.step.resize(cnt);
  }
  size_t w = 0;
  long long ev = 0;
  const int gb = net->gid_bits;
  const long long* od = net->outdeg.data();
  const unsigned long long* hr = net->h_rec;
  int32_t* og = blk.gid.data();
  int32_t* os = blk.step.data();
  if (net->all_recorded) {
    // every record is kept: a straight decode, split over host threads when
    // the chunk is large (the decode must finish inside the next chunk)
    const long long c0 = j.c0;
    auto decode = [&](size_t lo, size_t hi, long long* evp) {
      long long e = 0;
      for (size_t i = lo; i < hi; ++i) {
        const unsigned long long k = hr[i];
        const int g = (int)(k & gmask);
        e += od[g];
        og[i] = g;
        os[i] = (int)(c0 + (long long)(k >> gb));
      }
      *evp = e;
    };
    const unsigned hc = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const unsigned nth = cnt >= (1ull << 18) ? hc : 1u;
    std::vector<long long> evt(nth, 0);
    std::vector<std::thread> pool;
    const size_t per = (size_t)((cnt + nth - 1) / nth);
    for (unsigned t = 1; t < nth; ++t) {
      const size_t lo = std::min<size_t>(cnt, per * t), hi = std::min<size_t>(cnt, per * (t + 1));
      pool.emplace_back(decode, lo, hi, &evt[t]);
    }
    decode(0, std::min<size_t>(cnt, per), &evt[0]);
    for (std::thread& th : pool) th.join();
    for (long long e : evt) ev += e;
    w = cnt;
  } else {
    for (unsigned long long i = 0; i < cnt; ++i) {
      const unsigned long long k = hr[i];
      const int g = (int)(k & gmask);
      ev += od[g];
      if (net->pop_record[net->gid_pop[g]]) {
        og[w] = g;
        os[w] = (int)(j.c0 + (long long)(k >> gb));
        ++w;
      }
    }
  }
  blk.n = w;
  net->n_out += (long long)w;
  net->n_events += ev;
  // gather: each spike's bucket entry and segment bounds, one edge word per
  // edge (2 B, or 4 B when the words need 32 bits); 8 B record
  net->b_deliver += 8.0 * (double)cnt + (net->narrow ? 2.0 : 4.0) * (double)ev;
  net->b_update += 8.0 * (double)cnt;
  for (bs_net::WTap& t : net->wtaps)
    if (t.E > 0 && j.chunk > 0)
      HIPCHK(hipMemcpy(t.host.data() + (size_t)j.done * t.E, t.d_traj, sizeof(double) * (size_t)j.chunk * t.E,
                       hipMemcpyDeviceToHost));
  const int ntaps = (int)net->vm_tap_col.size();
  if (vm_out && ntaps > 0) {
    const double* hv;
    std::vector<double> hvm;
    if (pin_vm) {
      hv = net->h_vm_p + (size_t)j.par * net->vm_cap_steps * net->n_vm_dev;
    } else {
      hvm.resize((size_t)j.chunk * std::max(1, net->n_vm_dev));
      HIPCHK(hipMemcpy(hvm.data(), net->d_vm, 8 * hvm.size(), hipMemcpyDeviceToHost));
      hv = hvm.data();
    }
    for (long long t = 0; t < j.chunk; ++t)
      for (int k = 0; k < ntaps; ++k)
        vm_out[(j.done + t) * ntaps + k] = hv[(size_t)t * net->n_vm_dev + net->vm_tap_col[k]];
  }
  return BS_OK;
}

// The engine's sine (sin_p) reduces its argument by pi/2 in three parts whose
// products with the quadrant count are exact only while |argument| < 2^20 *
// pi/2 (about 1.6e6 rad); a run whose ac_generator arguments omega * n * dt +
// phi could leave that range is refused rather than evaluated inexactly.
static int check_ac_range(bs_net* net, long long n_end) {
  for (const auto& a : net->ac_args) {
    const double last = std::min((double)n_end, a[4]);
    if (last <= a[3]) continue;  // never active in this run
    const double arg = std::fabs(a[0]) * last * a[2] + std::fabs(a[1]);
    if (!(arg < 1.6e6))
      return fail(BS_EINVAL, "ac_generator: omega * t + phi reaches " + std::to_string(arg) +
                                 " rad in this run, beyond the 1.6e6 rad the engine's sine reduces exactly");
  }
  return BS_OK;
}

int bs_net_run(bs_net* net, int64_t n_steps, double* vm_out) {
  if (!net) return fail(BS_EINVAL, "bs_net_run: null handle");
  if (n_steps < 0) return fail(BS_EINVAL, "bs_net_run: n_steps must be >= 0");
  if (net->step + n_steps >= (1ll << 31)) return fail(BS_EINVAL, "bs_net_run: step counter overflow");
  {
    int rc = check_ac_range(net, net->step + n_steps);
    if (rc != BS_OK) return rc;
  }
  HIPCHK(hipSetDevice(net->device));
  net->n_out_blocks = 0;
  net->n_out = 0;
  net->t_deliver = net->t_update = net->t_total = 0;
  net->n_deliver = net->n_update = 0;
  net->b_deliver = net->b_update = 0;
  net->n_events = 0;
  const int ntaps = (int)net->vm_tap_col.size();
  // = K * W; record_weight taps bound it by their device trajectory rows
  const long long chunk_max = net->wtaps.empty() ? net->vm_cap_steps : std::min(net->vm_cap_steps, net->wtap_rows);
  for (bs_net::WTap& t : net->wtaps) t.host.assign((size_t)n_steps * t.E, 0.0);
  net->wtap_steps = n_steps;
  // Pipelined drain: chunk c + 1 is queued before chunk c is drained, so the
  // device never waits on the host's sort, copy and decode. Timed runs keep
  // one chunk in flight (their per-window events are read at each drain), and
  // so do runs whose V_m rows would need more than 256 MiB of pinned memory.
  const size_t vm_bytes = sizeof(double) * (size_t)chunk_max * std::max(1, net->n_vm_dev);
  const bool want_vm = vm_out && ntaps > 0;
  const bool pipelined = !net->timing && net->wtaps.empty() && n_steps > chunk_max &&
                         (!want_vm || vm_bytes <= (256u << 20));
  const bool pin_vm = pipelined && want_vm;
  if (pin_vm && net->h_vm_cap < 2 * vm_bytes) {
    if (net->h_vm_p) HIPCHK(hipHostFree(net->h_vm_p));
    net->h_vm_p = nullptr;
    net->h_vm_cap = 0;
    HIPCHK(hipHostMalloc((void**)&net->h_vm_p, 2 * vm_bytes, hipHostMallocDefault));
    net->h_vm_cap = 2 * vm_bytes;
  }
  hipEvent_t ev_all0, ev_all1;
  HIPCHK(hipEventCreate(&ev_all0));
  HIPCHK(hipEventCreate(&ev_all1));
  std::vector<hipEvent_t> evs;
  HIPCHK(hipEventRecord(ev_all0, net->stream));
  if (net->n_prod > 0 && n_steps > 0) {
    launch_pois_fill(net, (int)net->step);
    HIPCHK(hipGetLastError());
  }
  // Algorithmic update bytes: once per window every neuron reads and writes
  // its state variables and its int32 refractory counter (the window's steps
  // run with the state in registers); every step writes 8 B per V_m sample.
  double win_bytes = 0.0;
  for (const DevPop& p : net->pops) win_bytes += (double)p.n * (2.0 * p.n_state * net->tsize + 8.0);
  const double step_bytes = 8.0 * net->n_vm_dev;
  long long done = 0;
  int rc = BS_OK;
  int par = 0;
  ChunkJob pending;
  bool have_pending = false;
  while (rc == BS_OK && (done < n_steps || have_pending)) {
    ChunkJob cur;
    bool queued = false;
    if (done < n_steps) {
      cur.c0 = net->step;
      cur.chunk = std::min<long long>(chunk_max, n_steps - done);
      cur.done = done;
      cur.par = par;
      const int Wr = net->wtaps.empty() ? net->W : 1;
      cur.nwin = (int)((cur.chunk + Wr - 1) / Wr);
      rc = enqueue_chunk(net, cur, pin_vm, evs, win_bytes, step_bytes);
      if (rc != BS_OK) break;
      queued = true;
      net->step += cur.chunk;
      done += cur.chunk;
      par ^= 1;
    }
    if (have_pending) {
      rc = drain_chunk(net, pending, pin_vm, vm_out, evs);
      have_pending = false;
    }
    if (queued) {
      if (pipelined) {
        pending = cur;
        have_pending = true;
      } else if (rc == BS_OK) {
        rc = drain_chunk(net, cur, pin_vm, vm_out, evs);
      }
    }
  }
  // an error leaves at most one chunk queued: let it finish before returning
  (void)hipStreamSynchronize(net->stream);
  HIPCHK(hipEventRecord(ev_all1, net->stream));
  HIPCHK(hipEventSynchronize(ev_all1));
  float tot = 0;
  HIPCHK(hipEventElapsedTime(&tot, ev_all0, ev_all1));
  net->t_total = tot;
  net->n_deliver = net->n_update;
  net->t_deliver = net->t_update;
  for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ev_all0);
  (void)hipEventDestroy(ev_all1);
  for (void* p : net->h_rec_old) (void)hipHostFree(p);  // the device is idle here
  net->h_rec_old.clear();
  return rc;
}

// ---- device-resident, stream-ordered run ---------------------------------
__global__ void k_run_base(long long* base, long long b, int* prev, int p) {
  base[0] = b;
  prev[0] = p;
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, long long n, uint32_t v) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}
__global__ void k_zero_u64(unsigned long long* p) { p[0] = 0ull; }
// The chunk's spike records (key = step - chunk start << gid_bits | gid) as
// raster bits of rows row0 + (step - chunk start).
__global__ void k_rec_raster(const unsigned long long* __restrict__ rec, const unsigned long long* __restrict__ cnt,
                             long long cap, int gid_bits, long long row0, long long words,
                             uint32_t* __restrict__ raster) {
  const long long n = (long long)min((unsigned long long)cap, *cnt);
  const unsigned long long gmask = (1ull << gid_bits) - 1ull;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    const unsigned long long key = rec[k];
    const long long g = (long long)(key & gmask), row = row0 + (long long)(key >> gid_bits);
    atomicOr(&raster[row * words + (g >> 5)], 1u << (g & 31));
  }
}
// V_m rows of the chunk in caller tap order.
__global__ void k_vm_taps(const double* __restrict__ vbuf, int n_vm_dev, const int* __restrict__ tap_col, int ntaps,
                          long long rows, long long row0, double* __restrict__ out) {
  const long long n = rows * ntaps;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    const long long t = k / ntaps;
    const int c = (int)(k - t * ntaps);
    out[(row0 + t) * ntaps + c] = vbuf[t * n_vm_dev + tap_col[c]];
  }
}
__global__ void k_run_status(const int* __restrict__ flags, int32_t* __restrict__ status) {
  const int f = flags[0];
  status[0] = (f & (FLAG_OVERFLOW | FLAG_COUNT | FLAG_PLASTIC)) ? BS_ECAPACITY : ((f & FLAG_UNSTABLE) ? BS_ENUMERIC : BS_OK);
}

int bs_net_run_device(bs_net* net, int64_t n_steps, void* stream, uint32_t* raster, double* vm, int32_t* status) {
  if (!net) return fail(BS_EINVAL, "bs_net_run_device: null handle");
  if (n_steps < 0) return fail(BS_EINVAL, "bs_net_run_device: n_steps must be >= 0");
  if (n_steps > 0 && !raster) return fail(BS_EINVAL, "bs_net_run_device: null raster");
  if (net->step + n_steps >= (1ll << 31)) return fail(BS_EINVAL, "bs_net_run_device: step counter overflow");
  {
    int rc = check_ac_range(net, net->step + n_steps);
    if (rc != BS_OK) return rc;
  }
  HIPCHK(hipSetDevice(net->device));
  hipStream_t cs = (hipStream_t)stream;
  hipStream_t st = net->stream;
  HIPCHK(hipEventRecord(net->ev_join[0], cs));
  HIPCHK(hipStreamWaitEvent(st, net->ev_join[0], 0));
  const long long words = (net->Ntot + 31) / 32;
  const int ntaps = (int)net->vm_tap_col.size();
  auto grid = [](long long n) { return dim3((unsigned)std::max<long long>(1, std::min<long long>(1024, (n + 255) / 256))); };
  if (n_steps > 0) {
    hipLaunchKernelGGL(k_fill_u32, grid(n_steps * words), dim3(256), 0, st, raster, n_steps * words, 0u);
    if (net->n_prod > 0) launch_pois_fill(net, (int)net->step);
  }
  const long long chunk_max = net->vm_cap_steps;
  long long done = 0;
  int par = 0;
  while (done < n_steps) {
    const long long c0 = net->step;
    const long long chunk = std::min<long long>(chunk_max, n_steps - done);
    const int nwin = (int)((chunk + net->W - 1) / net->W);
    net->rec_par = par;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, net->d_rec_count + par);
    if (chunk == chunk_max) {  // a full chunk: replay its graph from this base step
      int rc = chunk_graph(net, par, nwin);
      if (rc != BS_OK) return rc;
      net->rec_par = par;
      hipLaunchKernelGGL(k_run_base, dim3(1), dim3(1), 0, st, net->d_base, (long long)c0, net->d_prevlen,
                         net->last_win);
      HIPCHK(hipGraphLaunch(net->graph[par], st));
      net->last_win = net->W;
    } else {
      for (int w = 0; w < nwin; ++w) {
        const long long n0 = c0 + (long long)w * net->W;
        const int Wc = (int)std::min<long long>(net->W, c0 + chunk - n0);
        if (net->tsize == 8)
          launch_window<double>(net, (int)n0, Wc, c0, net->last_win);
        else
          launch_window<float>(net, (int)n0, Wc, c0, net->last_win);
        net->last_win = Wc;
      }
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_rec_raster, grid(net->rec_cap), dim3(256), 0, st, (const unsigned long long*)net->d_rec[par],
                       (const unsigned long long*)(net->d_rec_count + par), net->rec_cap, net->gid_bits, done, words,
                       raster);
    if (vm && ntaps > 0)
      hipLaunchKernelGGL(k_vm_taps, grid(chunk * ntaps), dim3(256), 0, st, (const double*)net->d_vm, net->n_vm_dev,
                         (const int*)net->d_tap_col, ntaps, chunk, done, vm);
    HIPCHK(hipGetLastError());
    net->step += chunk;
    done += chunk;
    par ^= 1;
  }
  if (status) hipLaunchKernelGGL(k_run_status, dim3(1), dim3(1), 0, st, (const int*)net->d_flags, status);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(net->ev_join[1], st));
  HIPCHK(hipStreamWaitEvent(cs, net->ev_join[1], 0));
  net->n_out_blocks = 0;  // this run's spikes are in the caller's raster
  net->n_out = 0;
  return BS_OK;
}

// ---- whole-engine checkpoint -------------------------------------------------
namespace {
struct StateHdr {
  char magic[8];
  int abi, dtype, n_pops, Ls, Lr, NCW, NFX, W, n_stdp, last_win;
  long long Ntot, step, payload;
};
struct Part {
  void* p;
  size_t bytes;
};
std::vector<Part> state_parts(bs_net* net) {
  const size_t N = (size_t)net->Ntot;
  std::vector<Part> v;
  v.push_back({net->d_state, (size_t)BS_NSTATE * N * net->tsize});
  v.push_back({net->d_ref, 4 * N});
  v.push_back({net->d_sb, 4 * (size_t)net->Ls * N});
  v.push_back({net->d_sb_count, 4 * (size_t)net->Ls * net->pops.size()});
  v.push_back({net->d_ring_cnt, 4 * (size_t)net->Lr * net->NCW * N});
  if (net->NFX > 0) v.push_back({net->d_ring_fx, 8 * (size_t)net->Lr * net->NFX * N});
  if (net->d_spk_hist) v.push_back({net->d_spk_hist, (size_t)net->Ls * N});
  for (size_t q = 0; q < net->stdp.size(); ++q) {
    const StdpDev& S = net->stdp[q];
    v.push_back({S.w, 8 * net->w0[q].size()});
    v.push_back({S.U, 8 * (size_t)S.Lu * S.n_pre});
    v.push_back({S.V, 8 * (size_t)S.n_post});
    v.push_back({S.la, 4 * net->w0[q].size()});
  }
  return v;
}
StateHdr state_hdr(bs_net* net) {
  StateHdr h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "BSNETST", 8);
  h.abi = BS_ABI_VERSION;
  h.dtype = net->dtype;
  h.n_pops = (int)net->pops.size();
  h.Ls = net->Ls;
  h.Lr = net->Lr;
  h.NCW = net->NCW;
  h.NFX = net->NFX;
  h.W = net->W;
  h.n_stdp = (int)net->stdp.size();
  h.last_win = net->last_win;
  h.Ntot = net->Ntot;
  h.step = net->step;
  long long pay = 0;
  for (const Part& p : state_parts(net)) pay += (long long)p.bytes;
  h.payload = pay;
  return h;
}
}  // namespace

int bs_net_state_bytes(bs_net* net, int64_t* bytes) {
  if (!net || !bytes) return fail(BS_EINVAL, "bs_net_state_bytes: null argument");
  *bytes = (int64_t)sizeof(StateHdr) + state_hdr(net).payload;
  return BS_OK;
}

int bs_net_save_state(bs_net* net, void* buf, int64_t cap) {
  if (!net || !buf) return fail(BS_EINVAL, "bs_net_save_state: null argument");
  const StateHdr h = state_hdr(net);
  if (cap < (int64_t)sizeof(StateHdr) + h.payload) return fail(BS_ECAPACITY, "bs_net_save_state: buffer too small");
  HIPCHK(hipSetDevice(net->device));
  HIPCHK(hipStreamSynchronize(net->stream));
  unsigned char* o = (unsigned char*)buf;
  memcpy(o, &h, sizeof(h));
  o += sizeof(h);
  for (const Part& p : state_parts(net)) {
    if (p.bytes) HIPCHK(hipMemcpy(o, p.p, p.bytes, hipMemcpyDeviceToHost));
    o += p.bytes;
  }
  return BS_OK;
}

int bs_net_load_state(bs_net* net, const void* buf, int64_t size) {
  if (!net || !buf) return fail(BS_EINVAL, "bs_net_load_state: null argument");
  if (size < (int64_t)sizeof(StateHdr)) return fail(BS_EINVAL, "bs_net_load_state: not a snapshot");
  StateHdr h;
  memcpy(&h, buf, sizeof(h));
  StateHdr mine = state_hdr(net);
  if (memcmp(h.magic, "BSNETST", 8) != 0 || h.abi != BS_ABI_VERSION)
    return fail(BS_EINVAL, "bs_net_load_state: not a snapshot of this library version");
  if (h.dtype != mine.dtype || h.n_pops != mine.n_pops || h.Ls != mine.Ls || h.Lr != mine.Lr || h.NCW != mine.NCW ||
      h.NFX != mine.NFX || h.W != mine.W || h.n_stdp != mine.n_stdp || h.Ntot != mine.Ntot ||
      h.payload != mine.payload || size != (int64_t)sizeof(StateHdr) + h.payload)
    return fail(BS_EINVAL, "bs_net_load_state: the snapshot belongs to a different network");
  if (h.step < 0 || h.last_win < 0 || h.last_win > net->W)
    return fail(BS_EINVAL, "bs_net_load_state: corrupt snapshot header");
  HIPCHK(hipSetDevice(net->device));
  HIPCHK(hipStreamSynchronize(net->stream));
  const unsigned char* o = (const unsigned char*)buf + sizeof(h);
  for (const Part& p : state_parts(net)) {
    if (p.bytes) HIPCHK(hipMemcpy(p.p, o, p.bytes, hipMemcpyHostToDevice));
    o += p.bytes;
  }
  net->step = h.step;
  net->last_win = h.last_win;
  net->n_out_blocks = 0;
  net->n_out = 0;
  return BS_OK;
}

int bs_net_set_state(bs_net* net, int32_t pop, int32_t var, const double* in) {
  if (!net || !in) return fail(BS_EINVAL, "bs_net_set_state: null argument");
  if (pop < 0 || pop >= (int)net->pops.size()) return fail(BS_EINVAL, "bs_net_set_state: bad population");
  if (var < -1 || var >= BS_NSTATE) return fail(BS_EINVAL, "bs_net_set_state: bad variable");
  const DevPop& p = net->pops[pop];
  HIPCHK(hipSetDevice(net->device));
  HIPCHK(hipStreamSynchronize(net->stream));
  if (p.n == 0) return BS_OK;
  if (var == -1) {
    std::vector<int> r(p.n);
    for (int i = 0; i < p.n; ++i) {
      if (!(in[i] >= 0.0 && in[i] < 2147483648.0)) return fail(BS_EINVAL, "bs_net_set_state: bad refractory count");
      r[i] = (int)in[i];
    }
    HIPCHK(hipMemcpy(net->d_ref + p.gid0, r.data(), 4 * (size_t)p.n, hipMemcpyHostToDevice));
    return BS_OK;
  }
  const size_t off = (size_t)var * net->Ntot + p.gid0;
  if (net->tsize == 8) {
    HIPCHK(hipMemcpy((double*)net->d_state + off, in, 8 * (size_t)p.n, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(in, in + p.n);
    HIPCHK(hipMemcpy((float*)net->d_state + off, f.data(), 4 * (size_t)p.n, hipMemcpyHostToDevice));
  }
  return BS_OK;
}

int bs_net_set_drive(bs_net* net, int32_t channel, int64_t step0, int64_t n_steps, int32_t n_gen,
                     const int32_t* counts) {
  if (!net) return fail(BS_EINVAL, "bs_net_set_drive: null handle");
  if (channel < 0 || channel >= (int)net->chans.size() || net->chans[channel].kind != BS_CH_POISSON)
    return fail(BS_EINVAL, "bs_net_set_drive: channel is not a Poisson channel");
  DevChan& dc = net->chans[channel];
  if (step0 < 0 || n_steps < 0 || step0 + n_steps >= (1ll << 31))
    return fail(BS_EINVAL, "bs_net_set_drive: step range out of bounds");
  if (n_steps > 0 && (!counts || n_gen < dc.n_gen))
    return fail(BS_EINVAL, "bs_net_set_drive: counts must hold n_steps rows of at least " +
                               std::to_string(dc.n_gen) + " generator elements");
  const int w = std::max(1, dc.n_gen);
  std::vector<int32_t> rows((size_t)n_steps * w, 0);
  for (long long t = 0; t < n_steps; ++t)
    for (int e = 0; e < dc.n_gen; ++e) {
      const int32_t v = counts[(size_t)t * n_gen + e];
      if (v < 0 || v > 65535) return fail(BS_EINVAL, "bs_net_set_drive: counts must be in [0, 65535]");
      rows[(size_t)t * w + e] = v;
    }
  HIPCHK(hipSetDevice(net->device));
  HIPCHK(hipStreamSynchronize(net->stream));
  if (net->drive.empty()) net->drive.assign(net->chans.size(), nullptr);
  int32_t* old = net->drive[channel];
  int32_t* dd = nullptr;
  if (n_steps > 0) {
    HIPCHK(hipMalloc(&dd, sizeof(int32_t) * rows.size()));
    HIPCHK(hipMemcpy(dd, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice));
  }
  dc.drive = dd;
  dc.d0 = step0;
  dc.dn = n_steps;
  dc.dw = w;
  net->drive[channel] = dd;
  // the channel record lives in the channel array and in its population's blob
  HIPCHK(hipMemcpy(net->d_chans + channel, &dc, sizeof(DevChan), hipMemcpyHostToDevice));
  for (size_t p = 0; p < net->pops.size(); ++p) {
    const DevPop& dp = net->pops[p];
    if (channel < dp.ch_begin || channel >= dp.ch_begin + dp.ch_count) continue;
    if (p < net->pop_blob.size() && net->pop_blob[p] >= 0) {
      const size_t o = (size_t)net->pop_blob[p] + blob_layout(dp.ch_count, dp.gp_count, (size_t)dp.tab_bytes).ch +
                       sizeof(DevChan) * (size_t)(channel - dp.ch_begin);
      memcpy(&net->h_blob[o], &dc, sizeof(DevChan));
      HIPCHK(hipMemcpy(net->d_blob + o, &dc, sizeof(DevChan), hipMemcpyHostToDevice));
    }
  }
  if (old) HIPCHK(hipFree(old));
  return BS_OK;
}

int bs_net_set_weight_taps(bs_net* net, int32_t n, const int32_t* projs) {
  if (!net || n < 0 || (n > 0 && !projs)) return fail(BS_EINVAL, "bs_net_set_weight_taps: bad arguments");
  for (int k = 0; k < n; ++k)
    if (projs[k] < 0 || projs[k] >= (int)net->proj_stdp.size() || net->proj_stdp[projs[k]] < 0)
      return fail(BS_EINVAL, "bs_net_set_weight_taps: projection is not plastic");
  HIPCHK(hipSetDevice(net->device));
  HIPCHK(hipStreamSynchronize(net->stream));
  for (auto& t : net->wtaps) {
    (void)hipFree(t.d_ref);
    (void)hipFree(t.d_traj);
  }
  net->wtaps.clear();
  net->wtap_steps = 0;
  long long Esum = 0;
  for (int k = 0; k < n; ++k) Esum += (long long)net->w0[net->proj_stdp[projs[k]]].size();
  // trajectory rows per chunk: 256 MiB of device memory over all taps
  net->wtap_rows = std::max<long long>(1, std::min<long long>(net->vm_cap_steps, (32ll << 20) / std::max(1ll, Esum)));
  for (int k = 0; k < n; ++k) {
    bs_net::WTap t;
    t.q = net->proj_stdp[projs[k]];
    t.E = (long long)net->w0[t.q].size();
    net->wtaps.push_back(t);
    bs_net::WTap& b = net->wtaps.back();
    int rc = dmalloc(&b.d_ref, (size_t)std::max(1ll, b.E));
    if (rc) return rc;
    rc = dmalloc(&b.d_traj, (size_t)std::max(1ll, b.E * net->wtap_rows));
    if (rc) return rc;
    if (b.E) HIPCHK(hipMemcpy(b.d_ref, net->w_ref[b.q].data(), 8 * (size_t)b.E, hipMemcpyHostToDevice));
  }
  return BS_OK;
}

int bs_net_copy_weight_trace(bs_net* net, int32_t k, double* out) {
  if (!net || !out) return fail(BS_EINVAL, "bs_net_copy_weight_trace: null argument");
  if (k < 0 || k >= (int)net->wtaps.size()) return fail(BS_EINVAL, "bs_net_copy_weight_trace: no such tap");
  const std::vector<double>& h = net->wtaps[k].host;
  if (!h.empty()) memcpy(out, h.data(), sizeof(double) * h.size());
  return BS_OK;
}

int bs_net_window(bs_net* net, int32_t* W) {
  if (!net || !W) return fail(BS_EINVAL, "bs_net_window: null argument");
  *W = net->W;
  return BS_OK;
}

// A peer shard's window: its spikes become this handle's buckets of the
// remote populations for the steps of the window just run, which the next
// windows gather exactly like local buckets (the buckets of those steps were
// cleared by the window before).
int bs_net_inject_spikes(bs_net* net, int64_t n, const int32_t* gid, const int32_t* step) {
  if (!net) return fail(BS_EINVAL, "bs_net_inject_spikes: null handle");
  if (n < 0 || (n > 0 && (!gid || !step))) return fail(BS_EINVAL, "bs_net_inject_spikes: bad arguments");
  if (n == 0) return BS_OK;
  const long long lo = net->step - net->last_win, hi = net->step;
  // group by (step, population), keeping the given order inside a group
  std::vector<std::vector<uint32_t>> grp((size_t)std::max<long long>(1, hi - lo) * net->pops.size());
  for (int64_t i = 0; i < n; ++i) {
    const int g = gid[i];
    if (g < 0 || g >= net->Ntot) return fail(BS_EINVAL, "bs_net_inject_spikes: gid out of range");
    const int p = net->gid_pop[g];
    if (!net->pop_remote[p]) return fail(BS_EINVAL, "bs_net_inject_spikes: spike of a population local to this shard");
    if (step[i] < lo || step[i] >= hi)
      return fail(BS_EINVAL, "bs_net_inject_spikes: step outside the window this handle ran last");
    grp[(size_t)(step[i] - lo) * net->pops.size() + p].push_back((uint32_t)(g - net->pops[p].gid0));
  }
  HIPCHK(hipSetDevice(net->device));
  for (long long s = lo; s < hi; ++s)
    for (size_t p = 0; p < net->pops.size(); ++p) {
      const std::vector<uint32_t>& v = grp[(size_t)(s - lo) * net->pops.size() + p];
      if (v.empty()) continue;
      if ((long long)v.size() > net->pops[p].n) return fail(BS_EINVAL, "bs_net_inject_spikes: repeated spikes");
      const unsigned cnt = (unsigned)v.size();
      HIPCHK(hipMemcpy(net->d_sb + (size_t)(s % net->Ls) * net->Ntot + net->pops[p].gid0, v.data(),
                       sizeof(uint32_t) * v.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(net->d_sb_count + (size_t)(s % net->Ls) * net->pops.size() + p, &cnt, sizeof(unsigned),
                       hipMemcpyHostToDevice));
    }
  return BS_OK;
}

int bs_net_spike_count(bs_net* net, int64_t* n) {
  if (!net || !n) return fail(BS_EINVAL, "bs_net_spike_count: null argument");
  *n = (int64_t)net->n_out;
  return BS_OK;
}

int bs_net_copy_spikes(bs_net* net, int32_t* gid, int32_t* step, int64_t cap) {
  if (!net || !gid || !step) return fail(BS_EINVAL, "bs_net_copy_spikes: null argument");
  if (cap < (int64_t)net->n_out) return fail(BS_ECAPACITY, "bs_net_copy_spikes: cap too small");
  for (size_t b = 0; b < net->n_out_blocks; ++b) {
    const bs_net::SpikeBlock& blk = net->out_blocks[b];
    memcpy(gid, blk.gid.data(), 4 * blk.n);
    memcpy(step, blk.step.data(), 4 * blk.n);
    gid += blk.n;
    step += blk.n;
  }
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
  if (!net || !out) return fail(BS_EINVAL, "bs_net_get_weights: null argument");
  if (proj < 0 || proj >= (int)net->proj_stdp.size() || net->proj_stdp[proj] < 0)
    return fail(BS_EINVAL, "bs_net_get_weights: projection is not plastic");
  const int q = net->proj_stdp[proj];
  const size_t E = net->w0[q].size();
  std::vector<double> w(E);
  HIPCHK(hipStreamSynchronize(net->stream));
  if (E) HIPCHK(hipMemcpy(w.data(), net->stdp[q].w, 8 * E, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < E; ++k) out[k] = w[net->w_ref[q][k]];
  return BS_OK;
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
  else if (op == 3) {
    bool bad = false;
    const double q = div_rc(a[i], b[i], 1.0 / b[i], bad);
    o[i] = bad ? a[i] / b[i] : q;
  }
}
__global__ void k_dbg_rng(long long n, const uint32_t* k0, const uint32_t* k1, const uint32_t* c0,
                          const uint32_t* c1, uint32_t* o0, uint32_t* o1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  threefry2x32(k0[i], k1[i], c0[i], c1[i], o0[i], o1[i]);
}
__global__ void k_dbg_poisson(long long n, const uint32_t* seed, const uint32_t* step,
                              const uint32_t* elem, const double* cdf, const unsigned short* guide,
                              int len, int parts, uint32_t* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = poisson_inv(seed[i], step[i], elem[i], cdf, guide, len, parts);
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
  if (op < 0 || op > 3 || n < 0 || !a || !out) return fail(BS_EINVAL, "bs_debug_math: bad args");
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
  std::vector<double> cdf;
  std::vector<unsigned short> guide;
  int parts = 1;
  poisson_table(lam, cdf, guide, parts);
  uint32_t *a, *b, *c, *o;
  double* dc;
  unsigned short* dg;
  to_dev(&a, seed, n); to_dev(&b, step, n); to_dev(&c, elem, n); to_dev(&o, (const uint32_t*)nullptr, n);
  to_dev(&dc, cdf.data(), (long long)cdf.size()); to_dev(&dg, guide.data(), (long long)guide.size());
  hipLaunchKernelGGL(k_dbg_poisson, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (long long)n, a, b, c,
                     (const double*)dc, (const unsigned short*)dg, (int)cdf.size(), parts, o);
  hipError_t e = hipMemcpy(out, o, 4 * n, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); (void)hipFree(o); (void)hipFree(dc); (void)hipFree(dg);
  if (e != hipSuccess) return fail(BS_EHIP, hipGetErrorString(e));
  return BS_OK;
}

