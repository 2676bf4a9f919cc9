// oracle.cpp -- CPU restatement of the reference Simulator step (see oracle.h).
// TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's CPU leg.
//
// Every arithmetic expression below follows the reference source line it cites,
// in the same association order; the file is compiled with -ffp-contract=off
// so each operation is one IEEE-rounded op.
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#ifndef ORC_SOURCE_HASH
#define ORC_SOURCE_HASH "unknown"
#endif
extern "C" const char* orc_source_hash(void) { return ORC_SOURCE_HASH; }

static thread_local std::string g_err;
static int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}
extern "C" const char* orc_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------- primitives
// Threefry-2x32, 20 rounds (Salmon, Moraes, Dror, Shaw; SC'11, Random123):
// key schedule k2 = 0x1BD11BDA ^ k0 ^ k1; rotations 13,15,26,6,17,29,16,24;
// a key injection after every fourth round, with the injection counter added
// to the second word.
extern "C" void orc_threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1,
                                 uint32_t* o0, uint32_t* o1) {
  static const int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t key[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
  uint32_t x0 = c0 + key[0];
  uint32_t x1 = c1 + key[1];
  int round = 0;
  for (int inj = 1; inj <= 5; ++inj) {
    for (int q = 0; q < 4; ++q, ++round) {
      const int rr = rot[round % 8];
      x0 = x0 + x1;
      x1 = (x1 << rr) | (x1 >> (32 - rr));
      x1 = x1 ^ x0;
    }
    x0 = x0 + key[inj % 3];
    x1 = x1 + key[(inj + 1) % 3] + (uint32_t)inj;
  }
  *o0 = x0;
  *o1 = x1;
}

// Poisson sampling by inversion of a tabulated CDF -- the definition the
// device samples with. Table of Poisson(mu), mu = lam / parts with
// parts = max(1, ceil(lam / 32)): pmf_0 = exp(-mu), pmf_k = pmf_{k-1} * mu / k,
// cdf_k = cdf_{k-1} + pmf_k, stopping once k > mu and 1 - cdf_k < 1e-15 (at
// most 512 entries); guide[g] = min{k : cdf_k >= g / 64}. Part p takes uniform
// number p of threefry(key=(seed, step), ctr=(elem, p / 2)) (word p % 2),
// u = (x + 1/2) 2^-32, and adds min{k : u <= cdf_k} searched upward from
// guide[floor(64 u)] and capped at the last entry.
struct PoissonTable {
  std::vector<double> cdf;
  std::vector<int> guide;
  int parts = 1;
};

static PoissonTable make_table(double lam) {
  PoissonTable T;
  T.parts = std::max(1, (int)ceil(lam / 32.0));
  const double mu = lam / (double)T.parts;
  double pmf = exp(-mu);
  T.cdf.push_back(pmf);
  for (int k = 1; k < 512; ++k) {
    pmf = pmf * mu / (double)k;
    const double c = T.cdf.back() + pmf;
    T.cdf.push_back(c);
    if ((double)k > mu && 1.0 - c < 1e-15) break;
  }
  const int len = (int)T.cdf.size();
  for (int g = 0; g < 64; ++g) {
    const double t = (double)g / 64.0;
    int k = 0;
    while (k < len - 1 && T.cdf[k] < t) ++k;
    T.guide.push_back(k);
  }
  return T;
}

static uint32_t draw_poisson(const PoissonTable& T, uint32_t seed, uint32_t step, uint32_t elem) {
  uint32_t count = 0, w[2] = {0, 0};
  const int len = (int)T.cdf.size();
  for (int p = 0; p < T.parts; ++p) {
    if (p % 2 == 0) orc_threefry2x32(seed, step, elem, (uint32_t)(p / 2), &w[0], &w[1]);
    const double u = ((double)w[p % 2] + 0.5) * 2.3283064365386963e-10;
    int k = T.guide[(int)(u * 64.0)];
    while (k < len - 1 && u > T.cdf[k]) ++k;
    count += (uint32_t)k;
  }
  return count;
}

extern "C" uint32_t orc_poisson(uint32_t seed, uint32_t step, uint32_t elem, double lam) {
  return draw_poisson(make_table(lam), seed, step, elem);
}

// exp and log from first principles: Cody-Waite reduction by ln 2 (split
// into a 42-bit head and a tail), Taylor series of exp to degree 13, and
// log(m) = 2 artanh((m-1)/(m+1)) to degree 23, both series summed in the
// same Estrin grouping, op for op. Both are also what the device
// evaluates (aeif_cond_alpha's exponential current and the RKF45 step-size
// controller's fractional powers).
static const double kLn2Hi = 0x1.62e42fefa3800p-1;
static const double kLn2Lo = 0x1.ef35793c76730p-45;
static const double kInvLn2 = 0x1.71547652b82fep+0;

extern "C" double orc_exp(double x) {
  if (isnan(x)) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double k = rint(x * kInvLn2);
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  // Taylor series to r^13 as 1 + (r + r^2 (1/2 + r T)), T = sum_{n=3..13}
  // r^(n-3)/n! grouped in pairs, then by r^2, r^4, r^8 (Estrin's scheme).
  static const double f[14] = {1.0,           1.0,           0.5,
                               1.0 / 6.0,     1.0 / 24.0,    1.0 / 120.0,
                               1.0 / 720.0,   1.0 / 5040.0,  1.0 / 40320.0,
                               1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0,
                               1.0 / 479001600.0, 1.0 / 6227020800.0};
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  double a[5];
  for (int j = 0; j < 5; ++j) a[j] = f[3 + 2 * j] + f[4 + 2 * j] * r;
  const double b0 = a[0] + a[1] * r2, b1 = a[2] + a[3] * r2, b2 = a[4] + f[13] * r2;
  const double t = (b0 + b1 * r4) + b2 * r8;
  const double p = 1.0 + (r + r2 * (0.5 + r * t));
  return ldexp(p, (int)k);
}

// sin(x) of the ac_generator current (the reference calls jnp.sin,
// ac_generator.py:312-313). The same function the device evaluates: reduction
// by pi/2 in three parts (Cody & Waite; 33 + 33 + 53 bits, exact products for
// |k| < 2^20), then by quadrant the Taylor series of sin r to r^21 or of
// cos r to r^20, Horner in z = r^2 from the highest term, op for op.
static const double kPio2[3] = {0x1.921fb54400000p+0, 0x1.0b4611a600000p-34,
                                0x1.3198a2e037073p-69};
static const double kTwoOverPi = 0x1.45f306dc9c883p-1;

extern "C" double orc_sin(double x) {
  if (!isfinite(x)) return x - x;
  const double k = rint(x * kTwoOverPi);
  const double r = ((x - k * kPio2[0]) - k * kPio2[1]) - k * kPio2[2];
  const double z = r * r;
  // fact[n] = n! as a double literal's value: the series divide 1 by it
  static const double fact[22] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                                  362880.0, 3628800.0, 39916800.0, 479001600.0,
                                  6227020800.0, 87178291200.0, 1307674368000.0,
                                  20922789888000.0, 355687428096000.0, 6402373705728000.0,
                                  121645100408832000.0, 2432902008176640000.0,
                                  51090942171709440000.0};
  const long long q = (long long)k & 3;
  double v;
  if (q & 1) {  // cos r = 1 + z (-1/2! + z (1/4! - ...)), top term 1/20!
    double c = 1.0 / fact[20];
    for (int n = 18; n >= 2; n -= 2) c = ((n / 2) % 2 ? -1.0 : 1.0) / fact[n] + z * c;
    v = 1.0 + z * c;
  } else {      // sin r = r + r z (-1/3! + z (1/5! - ...)), top term 1/21!
    double c = 1.0 / fact[21];
    for (int n = 19; n >= 3; n -= 2) c = ((n / 2) % 2 ? -1.0 : 1.0) / fact[n] + z * c;
    v = r + (r * z) * c;
  }
  return (q & 2) ? -v : v;
}

extern "C" double orc_log(double x) {
  if (isnan(x) || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (isinf(x)) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.7071067811865476) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  // 1 + z/3 + z^2 Q(z), Q = sum_{j=0..9} z^j/(2j+5) grouped like exp's
  const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  double a[5];
  for (int j = 0; j < 5; ++j) a[j] = 1.0 / (double)(5 + 4 * j) + z * (1.0 / (double)(7 + 4 * j));
  const double qq = ((a[0] + a[1] * z2) + (a[2] + a[3] * z2) * z4) + a[4] * z8;
  const double q = 1.0 + z * (1.0 / 3.0 + z * qq);
  const double lm = 2.0 * s * q;
  const double ed = (double)e;
  return ed * kLn2Hi + (ed * kLn2Lo + lm);
}

extern "C" double orc_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
  return orc_exp(y * orc_log(x));
}

// ------------------------------------------------------------ the network
struct OrcProj {
  int pre_pop, post_pop, channel, kind;
  std::vector<int> row_ptr;          // CSR by pre (stable sort of the edge list)
  std::vector<int> post, delay;      // per CSR edge
  std::vector<double> w;             // per CSR edge weight (COUNT: channel weight)
  std::vector<int64_t> q;            // FIXED: rint(w * 2^F)
  // stdp_synapse (plastic projections)
  bool plastic = false;
  bs_stdp stdp;
  std::vector<int> pre;              // per CSR edge: pre local index
  std::vector<double> tr;            // per CSR edge: pre trace of its delayed spike train
  std::vector<double> V;             // per post neuron: post trace
};

struct orc_net {
  int dtype;
  double dt;
  std::vector<bs_population> pops;
  std::vector<std::vector<double>> params;  // per pop
  std::vector<bs_channel> chans;
  std::vector<std::vector<int>> gen_index;  // per channel (POISSON)
  std::vector<PoissonTable> ptab;           // per channel (POISSON)
  // per channel (POISSON): counts injected by orc_set_drive for generator
  // steps [drive0, drive0 + drive_n), rows of drive_w elements
  std::vector<std::vector<int32_t>> drive;
  std::vector<int64_t> drive0, drive_n;
  std::vector<int> drive_w;
  std::vector<std::vector<int64_t>> sched;  // per channel (SCHEDULE): plateau [start, stop) steps
  std::vector<std::vector<double>> sched_val;
  std::vector<OrcProj> projs;
  std::vector<int> chan_proj;               // channel -> projection (COUNT/FIXED)
  std::vector<long long> gid0;
  long long Ntot = 0;
  int Lh = 1;                               // holder history depth
  // state
  std::vector<std::vector<double>> st;      // per pop: [BS_NSTATE][n]
  std::vector<std::vector<int>> ref;
  std::vector<std::vector<std::vector<double>>> hist;  // per pop: [Lh][n] holder history
  std::vector<std::vector<std::vector<double>>> ghist; // per channel (POISSON): [Lh][n_post]
  std::vector<std::vector<std::vector<double>>> cring; // per channel (COUNT): arrival ring [Lh][n_post]
  std::vector<std::vector<std::vector<int64_t>>> qring;// per channel (FIXED): arrival ring [Lh][n_post]
  std::vector<int> chan_pop;                // channel -> post population
  long long step = 0;
  std::vector<int32_t> out_gid, out_step;
  std::vector<std::pair<int, int>> taps;    // (pop, local)
};

// ------------------------------------------------------- connectivity
// The device samplers' definition (bs_conn.hip), restated with plain integer
// arithmetic. Reference rules: sample_fixed_indegree with multapses
// (connectivity.py:101-136, candidates arange(0, j) ++ arange(j + 1, n_pre)
// under self-exclusion, :117-121) and sample_pairwise_bernoulli
// (connectivity.py:81-99: uniform < p over the pair grid, the diagonal masked,
// edges in row-major order). Random bits: Threefry-2x32-20 with key
// (seed & 0xFFFFFFFF, 0xE36F6E6E ^ (seed >> 32)).
static void conn_key(int64_t seed, uint32_t* k0, uint32_t* k1) {
  *k0 = (uint32_t)((uint64_t)seed & 0xFFFFFFFFull);
  *k1 = 0xE36F6E6Eu ^ (uint32_t)((uint64_t)seed >> 32);
}

// Post j, slot k: c = floor((x1 2^32 + x0) n_cand / 2^64), shifted past j
// under self-exclusion. Edge j * K + k.
extern "C" int orc_conn_fixed_indegree(int64_t seed, int32_t n_pre, int32_t n_post, int32_t K,
                                       int32_t exclude_self, int32_t* pre, int32_t* post) {
  if (n_pre < 0 || n_post < 0 || K < 0) return fail(BS_EINVAL, "orc_conn_fixed_indegree: negative size");
  const int64_t n_cand = exclude_self ? (int64_t)n_pre - 1 : (int64_t)n_pre;
  if ((int64_t)n_post * K > 0 && n_cand <= 0) return fail(BS_EINVAL, "no candidate presynaptic neurons");
  uint32_t k0, k1;
  conn_key(seed, &k0, &k1);
  for (int64_t j = 0; j < n_post; ++j)
    for (int64_t k = 0; k < K; ++k) {
      uint32_t x0, x1;
      orc_threefry2x32(k0, k1, (uint32_t)j, (uint32_t)k, &x0, &x1);
      const unsigned __int128 r = ((unsigned __int128)(((uint64_t)x1 << 32) | x0)) * (uint64_t)n_cand;
      uint32_t c = (uint32_t)(uint64_t)(r >> 64);
      if (exclude_self && c >= (uint32_t)j) c += 1;
      pre[j * K + k] = (int32_t)c;
      post[j * K + k] = (int32_t)j;
    }
  return BS_OK;
}

// Pair (i, j) is an edge iff x0 < ceil(p 2^32) (x0 * 2^-32 < p), i == j
// skipped under self-exclusion; row-major order. Writes at most cap edges and
// returns the number found (negative on a bad argument).
extern "C" int64_t orc_conn_bernoulli(int64_t seed, int32_t n_pre, int32_t n_post, double p,
                                      int32_t exclude_self, int32_t* pre, int32_t* post, int64_t cap) {
  if (n_pre < 0 || n_post < 0 || !(p >= 0.0)) return fail(BS_EINVAL, "orc_conn_bernoulli: bad sizes or p");
  uint64_t thr = 0;
  if (p >= 1.0) thr = 1ull << 32;
  else if (p > 0.0) thr = (uint64_t)ceil(p * 4294967296.0);
  uint32_t k0, k1;
  conn_key(seed, &k0, &k1);
  int64_t n = 0;
  for (int64_t i = 0; i < n_pre; ++i)
    for (int64_t j = 0; j < n_post; ++j) {
      if (exclude_self && i == j) continue;
      uint32_t x0, x1;
      orc_threefry2x32(k0, k1, (uint32_t)i, (uint32_t)j, &x0, &x1);
      if ((uint64_t)x0 < thr) {
        if (n < cap) {
          pre[n] = (int32_t)i;
          post[n] = (int32_t)j;
        }
        ++n;
      }
    }
  return n;
}

static int nstate_of(int model) {
  switch (model) {
    case BS_IAF_PSC_DELTA: return 2;
    case BS_IAF_PSC_EXP: return 5;
    case BS_IAF_PSC_ALPHA: return 6;
    case BS_AEIF_COND_ALPHA: return 8;
  }
  return -1;
}

extern "C" int orc_create(const bs_net_desc* d, orc_net** out) {
  if (!d || !out) return fail(BS_EINVAL, "orc_create: null argument");
  if (d->abi_version != BS_ABI_VERSION) return fail(BS_EINVAL, "orc_create: ABI mismatch");
  for (int q = 0; q < d->n_projs; ++q)
    if (d->projs[q].on_device) return fail(BS_EINVAL, "orc_create: the oracle reads host edge arrays only");
  for (int p = 0; p < d->n_pops; ++p)
    if (d->pops[p].remote) return fail(BS_EINVAL, "orc_create: the oracle runs whole networks (no remote populations)");
  orc_net* net = new orc_net();
  net->dtype = d->dtype;
  net->dt = d->dt_ms;
  net->pops.assign(d->pops, d->pops + d->n_pops);
  net->chans.assign(d->channels, d->channels + d->n_channels);
  net->chan_proj.assign(d->n_channels, -1);
  net->gen_index.resize(d->n_channels);
  net->ptab.resize(d->n_channels);
  net->sched.resize(d->n_channels);
  net->sched_val.resize(d->n_channels);
  long long g = 0;
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    if (nstate_of(sp.model) < 0 || !sp.state0) {
      delete net;
      return fail(BS_EINVAL, "orc_create: bad population");
    }
    net->params.emplace_back(sp.params, sp.params + BS_NPARAM);
    net->gid0.push_back(g);
    g += sp.n;
    for (int k = 0; k < sp.n_vm; ++k) net->taps.emplace_back(p, sp.vm_idx[k]);
    for (int c = sp.ch_begin; c < sp.ch_begin + sp.ch_count; ++c) {
      const bs_channel& ch = d->channels[c];
      if (ch.kind == BS_CH_SCHEDULE) {
        net->sched[c].assign(ch.sched, ch.sched + 2 * (size_t)std::max(0, ch.sched_len));
        net->sched_val[c].assign(ch.sched_val, ch.sched_val + std::max(0, ch.sched_len));
      }
      if (ch.kind == BS_CH_POISSON || ch.kind == BS_CH_CURRENT || ch.kind == BS_CH_AC || ch.kind == BS_CH_SCHEDULE) {
        std::vector<int> gi(sp.n);
        for (int i = 0; i < sp.n; ++i) gi[i] = ch.gen_index ? ch.gen_index[i] : i;
        net->gen_index[c] = gi;
        if (ch.kind == BS_CH_POISSON) net->ptab[c] = make_table(ch.lam);
      }
    }
  }
  net->Ntot = g;
  int dmax = 0;
  for (int c = 0; c < d->n_channels; ++c)
    if (d->channels[c].kind == BS_CH_POISSON)
      dmax = std::max(dmax, d->channels[c].delay_steps + (d->channels[c].delay_frac != 0.0 ? 1 : 0));
  for (int q = 0; q < d->n_projs; ++q) {
    const bs_projection& pj = d->projs[q];
    OrcProj P;
    P.pre_pop = pj.pre_pop;
    P.post_pop = pj.post_pop;
    P.channel = pj.channel;
    P.kind = d->channels[pj.channel].kind;
    const int npre = d->pops[pj.pre_pop].n;
    // csr_structure (connectivity.py:39-54): stable argsort by pre.
    std::vector<long long> order(pj.n_edges);
    for (long long e = 0; e < pj.n_edges; ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(),
                     [&](long long a, long long b) { return pj.pre[a] < pj.pre[b]; });
    P.row_ptr.assign(npre + 1, 0);
    for (long long e = 0; e < pj.n_edges; ++e) P.row_ptr[pj.pre[e] + 1]++;
    for (int i = 0; i < npre; ++i) P.row_ptr[i + 1] += P.row_ptr[i];
    const bs_channel& ch = d->channels[pj.channel];
    for (long long k = 0; k < pj.n_edges; ++k) {
      const long long e = order[k];
      P.post.push_back(pj.post[e]);
      const int dd = pj.delay ? pj.delay[e] : pj.delay_steps;
      P.delay.push_back(dd);
      dmax = std::max(dmax, dd);
      const double w = (P.kind == BS_CH_COUNT) ? ch.weight : pj.weight[e];
      P.w.push_back(w);
      if (P.kind == BS_CH_FIXED) P.q.push_back((int64_t)llrint(ldexp(w, ch.fixed_shift)));
      P.pre.push_back(pj.pre[e]);
    }
    if (pj.plastic) {
      P.plastic = true;
      P.stdp = pj.stdp;
      P.tr.assign(pj.n_edges, 0.0);
      P.V.assign(d->pops[pj.post_pop].n, 0.0);
    }
    net->chan_proj[pj.channel] = q;
    net->projs.push_back(std::move(P));
  }
  net->Lh = dmax + 2;
  *out = net;
  // initial state (init_all_states)
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    net->st.emplace_back(sp.state0, sp.state0 + (size_t)BS_NSTATE * sp.n);
    net->ref.emplace_back(sp.n, 0);
    net->hist.emplace_back(net->Lh, std::vector<double>(sp.n, 0.0));
  }
  net->ghist.resize(d->n_channels);
  net->drive.assign(d->n_channels, {});
  net->drive0.assign(d->n_channels, 0);
  net->drive_n.assign(d->n_channels, 0);
  net->drive_w.assign(d->n_channels, 0);
  net->cring.resize(d->n_channels);
  net->qring.resize(d->n_channels);
  net->chan_pop.assign(d->n_channels, 0);
  for (int p = 0; p < d->n_pops; ++p) {
    const bs_population& sp = d->pops[p];
    for (int c = sp.ch_begin; c < sp.ch_begin + sp.ch_count; ++c)
    {
      net->chan_pop[c] = p;
      if (d->channels[c].kind == BS_CH_POISSON)
        net->ghist[c].assign(net->Lh, std::vector<double>(sp.n, 0.0));
      else if (d->channels[c].kind == BS_CH_CURRENT || d->channels[c].kind == BS_CH_AC ||
               d->channels[c].kind == BS_CH_SCHEDULE)
        continue;  // no arrival ring: the current enters in its own step
      else if (d->channels[c].kind == BS_CH_COUNT)
        net->cring[c].assign(net->Lh, std::vector<double>(sp.n, 0.0));
      else
        net->qring[c].assign(net->Lh, std::vector<int64_t>(sp.n, 0));
    }
  }
  return BS_OK;
}

extern "C" int orc_destroy(orc_net* net) {
  delete net;
  return BS_OK;
}

// ----------------------------------------------------------- model updates
// iaf_psc_delta.update (_nest_neuron/iaf_psc_delta.py:478-521);
// i_total = sum_current_inputs(I_e + x): (I_e + 0) plus this step's current
// inputs in registration order (simulator.py:1737-1745).
template <typename T>
static bool upd_delta(const std::vector<double>& P, T& V, T& buf, int& r, T delta_v, T i_total) {
  const T E_L = (T)P[0], decay = (T)P[1];
  const T C_m = (T)P[32], tau_m = (T)P[33];
  const T last_v = V;
  const int ref_steps = r;
  T v_candidate = E_L + (last_v - E_L) * decay + (i_total / C_m) * tau_m * ((T)1 - decay);
  v_candidate = v_candidate + delta_v;
  const bool refractory_input = P[7] != 0.0;
  if (refractory_input) v_candidate = v_candidate + buf;
  if (P[5] > -1e300) v_candidate = std::max(v_candidate, (T)P[5]);
  const bool not_refractory = ref_steps == 0;
  const T v_post = not_refractory ? v_candidate : last_v;
  if (refractory_input) {
    const T refr_decay = (T)P[8 + std::min(ref_steps, 23)];  // exp(-ref_steps*dt/tau_m)
    buf = not_refractory ? (T)0 : (T)(buf + delta_v * refr_decay);
  }
  const int rs = not_refractory ? ref_steps : ref_steps - 1;
  const bool spike_cond = v_post >= (T)P[3];
  r = spike_cond ? (int)P[6] : rs;
  V = spike_cond ? (T)P[4] : v_post;
  return spike_cond;
}

// iaf_psc_exp.update (_nest_neuron/iaf_psc_exp.py:604-682), delta = 0
template <typename T>
static bool upd_exp(const std::vector<double>& P, T* s, int& r, T w_all, T cur) {
  const T E_L = (T)P[0];
  T V_rel = s[0] - E_L;
  const T theta = (T)P[1];
  T i_syn_ex = s[1], i_syn_in = s[2];
  const T i_0 = s[3], i_1 = s[4];
  const bool not_refractory = r == 0;
  const T V_candidate = V_rel * (T)P[3] + i_syn_ex * (T)P[4] + i_syn_in * (T)P[5] +
                        ((T)P[9] + i_0) * (T)P[6];
  V_rel = not_refractory ? V_candidate : V_rel;
  int rr = not_refractory ? r : r - 1;
  i_syn_ex = i_syn_ex * (T)P[7];
  i_syn_in = i_syn_in * (T)P[8];
  i_syn_ex = i_syn_ex + ((T)1 - (T)P[7]) * i_1;
  const T w_ex = w_all > (T)0 ? w_all : (T)0;
  const T w_in = w_all < (T)0 ? w_all : (T)0;
  i_syn_ex = i_syn_ex + w_ex;
  i_syn_in = i_syn_in + w_in;
  const bool spike_cond = V_rel >= theta;
  rr = spike_cond ? (int)P[10] : rr;
  V_rel = spike_cond ? (T)P[2] : V_rel;
  s[0] = V_rel + E_L;
  s[1] = i_syn_ex;
  s[2] = i_syn_in;
  s[3] = cur + (T)0;  // new i_0 = sum_current_inputs(x) + 0 pA, applied next step
  s[4] = (T)0;
  r = rr;
  return spike_cond;
}

// iaf_psc_alpha.update (_nest_neuron/iaf_psc_alpha.py:604-684)
template <typename T>
static bool upd_alpha(const std::vector<double>& P, T* s, int& r, T w_all, T cur) {
  const T w_ex = w_all > (T)0 ? w_all : (T)0;
  const T w_in = w_all < (T)0 ? w_all : (T)0;
  const T E_L = (T)P[0];
  T y3 = s[0] - E_L;
  T I_ex = s[1], I_in = s[2], dI_ex = s[3], dI_in = s[4];
  const T y0 = s[5];
  const bool not_refractory = r == 0;
  T y3c = (T)P[10] * (y0 + (T)P[17]) + (T)P[11] * dI_ex + (T)P[12] * I_ex + (T)P[13] * dI_in +
          (T)P[14] * I_in + (T)P[9] * y3 + y3;
  if (P[19] > -1e300) y3c = std::max(y3c, (T)P[19]);
  y3 = not_refractory ? y3c : y3;
  int rr = not_refractory ? r : r - 1;
  I_ex = (T)P[4] * dI_ex + (T)P[5] * I_ex;
  dI_ex = dI_ex * (T)P[3];
  dI_ex = dI_ex + (T)P[15] * w_ex;
  I_in = (T)P[7] * dI_in + (T)P[8] * I_in;
  dI_in = dI_in * (T)P[6];
  dI_in = dI_in + (T)P[16] * w_in;
  const bool spike_cond = y3 >= (T)P[1];
  rr = spike_cond ? (int)P[18] : rr;
  y3 = spike_cond ? (T)P[2] : y3;
  s[0] = y3 + E_L;
  s[1] = I_ex;
  s[2] = I_in;
  s[3] = dI_ex;
  s[4] = dI_in;
  s[5] = cur + (T)0;  // y0 for the next step: sum_current_inputs(x) + 0 pA
  r = rr;
  return spike_cond;
}

// aeif_cond_alpha: _vector_field (aeif_cond_alpha.py:554-592), _event_fn
// (:594-625), AdaptiveRungeKuttaStep.__call__ (_nest_base/utils.py:583-723)
// with the RKF45 tableau (utils.py:266-278) and atol = gsl_error_tol, rtol = 0.
namespace aeif {
enum { VPEAK = 0, VRESET, GL, CM, EEX, EIN, EL, DELTAT, TAUW, A, B, VTH, TAUEX, TAUIN, IE,
       ATOL, REFN, PSCEX, PSCIN, DT, MINH, VPDET, DTSAFE, MAXIT };
static const int NV = 6;  // V, dg_ex, g_ex, dg_in, g_in, w

static void vf(const std::vector<double>& P, const double* y, bool is_refractory, double i_stim,
               double* dy) {
  const double V = y[0], dg_ex = y[1], g_ex = y[2], dg_in = y[3], g_in = y[4], w = y[5];
  const double v_eff = is_refractory ? P[VRESET] : std::min(V, P[VPEAK]);
  const double i_syn_exc = g_ex * (v_eff - P[EEX]);
  const double i_syn_inh = g_in * (v_eff - P[EIN]);
  const double delta_t_safe = P[DELTAT] == 0.0 ? 1.0 : P[DELTAT];
  double exp_arg = (v_eff - P[VTH]) / delta_t_safe;
  exp_arg = std::min(std::max(exp_arg, -500.0), 500.0);
  const double i_spike = P[GL] * P[DELTAT] * orc_exp(exp_arg);
  const double dV_raw =
      (-P[GL] * (v_eff - P[EL]) + i_spike - i_syn_exc - i_syn_inh - w + P[IE] + i_stim) / P[CM];
  dy[0] = is_refractory ? 0.0 : dV_raw;
  dy[1] = -dg_ex / P[TAUEX];
  dy[2] = dg_ex - g_ex / P[TAUEX];
  dy[3] = -dg_in / P[TAUIN];
  dy[4] = dg_in - g_in / P[TAUIN];
  dy[5] = (P[A] * (v_eff - P[EL]) - w) / P[TAUW];
}

// Butcher tableau entries as the Python float expressions of utils.py:266-278.
static const double Ac[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 4, 0, 0, 0, 0},
    {3.0 / 32, 9.0 / 32, 0, 0, 0},
    {1932.0 / 2197, -7200.0 / 2197, 7296.0 / 2197, 0, 0},
    {439.0 / 216, -8.0, 3680.0 / 513, -845.0 / 4104, 0},
    {-8.0 / 27, 2.0, -3544.0 / 2565, 1859.0 / 4104, -11.0 / 40}};
static const double Bc[6] = {16.0 / 135, 0.0, 6656.0 / 12825, 28561.0 / 56430, -9.0 / 50, 2.0 / 55};
static const double Bh[6] = {25.0 / 216, 0.0, 1408.0 / 2565, 2197.0 / 4104, -1.0 / 5, 0.0};

// _rk_weighted_sum: state + h * sum_{c != 0} c_j k_j (left fold).
static void wsum(const double* s, double h, const double* c, int m, double k[][NV], double* out) {
  for (int v = 0; v < NV; ++v) {
    bool first = true;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) {
      if (c[j] == 0.0) continue;
      if (first) {
        acc = c[j] * k[j][v];
        first = false;
      } else {
        acc = acc + c[j] * k[j][v];
      }
    }
    out[v] = first ? s[v] : s[v] + h * acc;
  }
}

static double nanmax(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  return a > b ? a : b;
}

static bool update(const std::vector<double>& P, double* st, int& r, double w_ex, double w_in,
                   double cur, bool* unstable) {
  double y[NV] = {st[0], st[1], st[2], st[3], st[4], st[5]};
  double h = st[6];
  const double i_stim = st[7];
  const double dt = P[DT], min_h = P[MINH], atol = P[ATOL];
  const int ref_count = (int)P[REFN];
  const double v_peak_detect = P[VPDET];
  double b_error[6];
  for (int j = 0; j < 6; ++j) b_error[j] = Bc[j] - Bh[j];
  bool spike_mask = false;
  h = std::max(h, min_h);
  double t_loc = 0.0;
  const int max_iters = (int)P[MAXIT];
  for (int n_iters = 0; t_loc < dt && n_iters < max_iters; ++n_iters) {
    h = std::max(min_h, std::min(h, dt - t_loc));
    const bool is_ref = r > 0;
    double k[6][NV], yi[NV];
    for (int i = 0; i < 6; ++i) {
      if (i == 0)
        for (int v = 0; v < NV; ++v) yi[v] = y[v];
      else
        wsum(y, h, Ac[i], i, k, yi);
      vf(P, yi, is_ref, i_stim, k[i]);
    }
    double y_high[NV], y_err[NV];
    wsum(y, h, Bc, 6, k, y_high);
    for (int v = 0; v < NV; ++v) {
      bool first = true;
      double acc = 0.0;
      for (int j = 0; j < 6; ++j) {
        if (b_error[j] == 0.0) continue;
        acc = first ? b_error[j] * k[j][v] : acc + b_error[j] * k[j][v];
        first = false;
      }
      y_err[v] = h * acc;
    }
    double err = 0.0;
    for (int v = 0; v < NV; ++v) {
      const double e = fabs(y_err[v]) / (atol + 0.0 * nanmax(fabs(y[v]), fabs(y_high[v])));
      err = v == 0 ? e : nanmax(err, e);
    }
    if (isnan(err)) err = INFINITY;
    const bool accept = (err <= 1.0) || (h <= min_h);
    const bool reject = !accept;
    if (accept) {
      for (int v = 0; v < NV; ++v) y[v] = y_high[v];
      t_loc = t_loc + h;
      if (y[0] < -1e3 || y[5] < -1e6 || y[5] > 1e6) *unstable = true;
      if (r > 0) y[0] = P[VRESET];
      const bool spike_now = (r <= 0) && (y[0] >= v_peak_detect);
      if (spike_now) {
        spike_mask = true;
        y[0] = P[VRESET];
        y[5] = y[5] + P[B];
        if (ref_count > 0) r = ref_count + 1;
      }
    }
    const double err_sg = std::max(err, 1e-30);
    const double inv_err = 1.0 / err_sg;
    const double fac_accept =
        err_sg <= 1e-30 ? 5.0 : std::min(std::max(0.9 * orc_pow(inv_err, 1.0 / 5), 0.2), 5.0);
    const double fac_reject = std::min(std::max(0.9 * orc_pow(inv_err, 1.0 / 4), 0.2), 1.0);
    if (accept) h = std::max(min_h, h * fac_accept);
    if (reject) h = std::max(min_h, h * fac_reject);
  }
  r = r > 0 ? r - 1 : r;
  y[1] = y[1] + P[PSCEX] * w_ex;
  y[3] = y[3] + P[PSCIN] * w_in;
  for (int v = 0; v < NV; ++v) st[v] = y[v];
  st[6] = h;
  st[7] = cur + 0.0;  // I_stim for the next step: sum_current_inputs(x) + 0 pA
  return spike_mask;
}
}  // namespace aeif

// One edge's stdp_synapse update (stdp_synapse.py:163-179): facilitation
// gated on the post spike with K+ = pre trace minus this step's pre spike,
// then depression gated on the pre spike with K- = post trace minus this
// step's post spike; each clamps to [0, Wmax] in Wmax-normalised form.
double orc_stdp_edge(double w, double pre_spike, double post_spike, double pre_trace,
                     double post_trace, const bs_stdp* S) {
  const double W = S->Wmax;
  const double kplus = pre_trace - pre_spike;
  const double kminus = post_trace - post_spike;
  if (post_spike > 0) {
    const double nw = w / W + S->lambda_ * orc_pow(1.0 - w / W, S->mu_plus) * kplus;
    w = nw < 1.0 ? nw * W : W;
  }
  if (pre_spike > 0) {
    const double nw = w / W - S->alpha * S->lambda_ * orc_pow(w / W, S->mu_minus) * kminus;
    w = nw > 0.0 ? nw * W : 0.0;
  }
  return w;
}

// stdp_synapse on EventPlasticProj for one step n (event_plastic.py:454-536,
// stdp_synapse.py:163-179): traces decay-then-add (the pre trace of each
// edge follows that edge's delayed spike train), facilitation where the post
// neuron fired in the previous step, then depression where the delayed pre
// spike arrives, then the new weights of the arriving edges are summed per
// post neuron (2^-F fixed point, the engine's order-free sum).
template <typename T>
static void stdp_step(orc_net* net, OrcProj& P, long long n, std::vector<T>& y) {
  const bs_stdp& S = P.stdp;
  const int Lh = net->Lh;
  const int F = net->chans[P.channel].fixed_shift;
  const std::vector<double>* post_prev =
      n >= 1 ? &net->hist[P.post_pop][(int)((n - 1) % Lh)] : nullptr;
  for (size_t j = 0; j < P.V.size(); ++j)
    P.V[j] = P.V[j] * S.decay_minus + (post_prev ? (*post_prev)[j] : 0.0);
  std::vector<int64_t> acc(y.size(), 0);
  for (size_t e = 0; e < P.w.size(); ++e) {
    const long long src = n - 1 - P.delay[e];
    const double x = src >= 0 ? net->hist[P.pre_pop][(int)(src % Lh)][P.pre[e]] : 0.0;
    P.tr[e] = P.tr[e] * S.decay_plus + x;
    const int j = P.post[e];
    const double post_fired = post_prev ? (*post_prev)[j] : 0.0;
    const double w = orc_stdp_edge(P.w[e], x, post_fired, P.tr[e], P.V[j], &S);
    if (x > 0) acc[j] += (int64_t)llrint(ldexp(w, F));
    P.w[e] = w;
  }
  for (size_t j = 0; j < y.size(); ++j) y[j] = (T)ldexp((double)acc[j], -F);
}

// -------------------------------------------------------------------- run
template <typename T>
static int run_t(orc_net* net, int64_t n_steps, double* vm_out) {
  const int np = (int)net->pops.size();
  const int Lh = net->Lh;
  const int ntaps = (int)net->taps.size();
  std::vector<std::vector<T>> yin(net->chans.size());  // per channel, this step's contribution
  for (int64_t t = 0; t < n_steps; ++t) {
    const long long n = net->step;
    const int hs = (int)(n % Lh);
    // ---- phase 1: projections read the holders of step n-1 through the delays
    for (size_t c = 0; c < net->chans.size(); ++c) {
      const bs_channel& ch = net->chans[c];
      const int post_pop = net->chan_pop[c];
      const int npost = net->pops[post_pop].n;
      yin[c].assign(npost, (T)0);
      if (ch.kind == BS_CH_POISSON) {
        // one_to_one: y = x_delayed * weight (event_proj.py:311-312); x is the
        // generator holder of step n-1 read d steps back.
        const long long src = n - 1 - ch.delay_steps;
        if (src < 0) continue;
        const std::vector<double>& x = net->ghist[c][(int)(src % Lh)];
        if (ch.delay_frac == 0.0) {
          for (int j = 0; j < npost; ++j) yin[c][j] = (T)x[j] * (T)ch.weight;
        } else {
          // InputDelay (delay.py:114-128) with a fractional delay/dt: frames
          // lo and lo + 1 mixed as x_lo * (1 - frac) + x_hi * frac, in the
          // network dtype; frames before step 0 hold zeros
          const T f = (T)ch.delay_frac;
          const std::vector<double>* xh = src >= 1 ? &net->ghist[c][(int)((src - 1) % Lh)] : nullptr;
          for (int j = 0; j < npost; ++j) {
            const T hi = xh ? (T)(*xh)[j] : (T)0;
            yin[c][j] = ((T)x[j] * ((T)1 - f) + hi * f) * (T)ch.weight;
          }
        }
      } else if (ch.kind == BS_CH_CURRENT) {
        // phase 1b: dc_generator current, amplitude * weight inside its window
        // [t_min_step, t_max_step), scattered onto its targets
        // (dc_generator.py:219-262, simulator.py:1737-1745)
        const bool active = ch.t_min_step <= n && n < ch.t_max_step;
        const std::vector<int>& gi = net->gen_index[c];
        for (int j = 0; j < npost; ++j) yin[c][j] = (active && gi[j] >= 0) ? (T)ch.weight : (T)0;
      } else if (ch.kind == BS_CH_SCHEDULE) {
        // step_current_generator (step_current_generator.py:269-335): the
        // plateau holding step n, by a plain scan of the schedule
        double v = 0.0;
        const std::vector<int64_t>& se = net->sched[c];
        for (size_t k = 0; k < net->sched_val[c].size(); ++k)
          if (se[2 * k] <= n && n < se[2 * k + 1]) v = net->sched_val[c][k];
        const std::vector<int>& gi = net->gen_index[c];
        for (int j = 0; j < npost; ++j) yin[c][j] = gi[j] >= 0 ? (T)v : (T)0;
      } else if (ch.kind == BS_CH_AC) {
        // phase 1b: ac_generator current amplitude * sin(omega t + phi) + offset
        // at t = n dt, times the connection weight, inside [t_min_step,
        // t_max_step) (ac_generator.py:300-325, simulator.py:1737-1745)
        const bool active = ch.t_min_step <= n && n < ch.t_max_step;
        const double t = (double)n * net->dt;
        const double I = (ch.ac[0] * orc_sin(ch.ac[1] * t + ch.ac[2]) + ch.ac[3]) * ch.weight;
        const std::vector<int>& gi = net->gen_index[c];
        for (int j = 0; j < npost; ++j) yin[c][j] = (active && gi[j] >= 0) ? (T)I : (T)0;
      }
    }
    // Static projections, in the kernel's DelayScatter form (delay_scatter.py:
    // 243-290): the spikes held for this step (emitted at n-1) are deposited
    // edge by edge, ascending pre then CSR order, into the arrival ring slot
    // (n + d_e); the slot of step n is then read. For a homogeneous delay this
    // is InputDelay followed by the CSR event matmul (delay.py:114-128,
    // projections.py:56-59), accumulated in the same sequential order.
    for (OrcProj& P : net->projs) {
      const int c = P.channel;
      std::vector<T>& y = yin[c];
      const int npre = net->pops[P.pre_pop].n;
      if (P.plastic) {
        stdp_step(net, P, n, y);
        continue;
      }
      if (n >= 1) {
        const std::vector<double>& x = net->hist[P.pre_pop][(int)((n - 1) % Lh)];
        for (int i = 0; i < npre; ++i) {
          if (x[i] == 0.0) continue;
          for (int e = P.row_ptr[i]; e < P.row_ptr[i + 1]; ++e) {
            const int slot = (int)((n + P.delay[e]) % Lh);
            if (P.kind == BS_CH_COUNT) {
              double& r = net->cring[c][slot][P.post[e]];
              r = (double)((T)r + (T)P.w[e]);
            } else {
              net->qring[c][slot][P.post[e]] += P.q[e];
            }
          }
        }
      }
      const int slot = (int)(n % Lh);
      if (P.kind == BS_CH_COUNT) {
        std::vector<double>& r = net->cring[c][slot];
        for (size_t j = 0; j < y.size(); ++j) {
          y[j] = (T)r[j];
          r[j] = 0.0;
        }
      } else {
        const int F = net->chans[c].fixed_shift;
        std::vector<int64_t>& r = net->qring[c][slot];
        for (size_t j = 0; j < y.size(); ++j) {
          y[j] = (T)ldexp((double)r[j], -F);
          r[j] = 0;
        }
      }
    }
    // ---- phase 2: generators draw, neurons update, holders capture
    for (size_t c = 0; c < net->chans.size(); ++c) {
      const bs_channel& ch = net->chans[c];
      if (ch.kind != BS_CH_POISSON) continue;
      std::vector<double>& h = net->ghist[c][hs];
      const std::vector<int>& gi = net->gen_index[c];
      const bool active = (ch.t_min_step < n) && (n <= ch.t_max_step) && ch.lam > 0.0;
      // poisson_generator.update (poisson_generator.py:501-576): counts of
      // this step, kept inside the activity window; an injected count
      // stands for the draw
      const bool inj = !net->drive[c].empty() && n >= net->drive0[c] && n < net->drive0[c] + net->drive_n[c];
      for (size_t j = 0; j < h.size(); ++j) {
        if (!active || gi[j] < 0) {
          h[j] = 0.0;
        } else if (inj) {
          h[j] = (double)net->drive[c][(size_t)(n - net->drive0[c]) * net->drive_w[c] + gi[j]];
        } else {
          h[j] = (double)draw_poisson(net->ptab[c], ch.seed, (uint32_t)n, (uint32_t)gi[j]);
        }
      }
    }
    for (int p = 0; p < np; ++p) {
      const bs_population& sp = net->pops[p];
      const std::vector<double>& P = net->params[p];
      std::vector<double>& S = net->st[p];
      std::vector<double>& H = net->hist[p][hs];
      for (int j = 0; j < sp.n; ++j) {
        T acc[2] = {(T)0, (T)0};  // sum_delta_inputs(zeros) per receptor
        // sum_current_inputs(init): I_e + x for iaf_psc_delta, x for the others (x = 0)
        T cur = sp.model == BS_IAF_PSC_DELTA ? (T)P[34] + (T)0 : (T)0;
        for (int c = sp.ch_begin; c < sp.ch_begin + sp.ch_count; ++c) {
          if (net->chans[c].kind == BS_CH_CURRENT || net->chans[c].kind == BS_CH_AC ||
              net->chans[c].kind == BS_CH_SCHEDULE)
            cur = cur + yin[c][j];
          else
            acc[net->chans[c].receptor] = acc[net->chans[c].receptor] + yin[c][j];
        }
        T s[BS_NSTATE];
        for (int k = 0; k < BS_NSTATE; ++k) s[k] = (T)S[(size_t)k * sp.n + j];
        int& r = net->ref[p][j];
        bool spk = false;
        if (sp.model == BS_IAF_PSC_DELTA) {
          spk = upd_delta<T>(P, s[0], s[1], r, acc[0], cur);
        } else if (sp.model == BS_IAF_PSC_EXP) {
          spk = upd_exp<T>(P, s, r, acc[0], cur);
        } else if (sp.model == BS_IAF_PSC_ALPHA) {
          spk = upd_alpha<T>(P, s, r, acc[0], cur);
        } else {
          double sd[BS_NSTATE];
          for (int k = 0; k < BS_NSTATE; ++k) sd[k] = (double)s[k];
          bool unstable = false;
          spk = aeif::update(P, sd, r, (double)acc[0], (double)acc[1], (double)cur, &unstable);
          if (unstable) return fail(BS_ENUMERIC, "Numerical instability in aeif_cond_alpha dynamics.");
          for (int k = 0; k < BS_NSTATE; ++k) s[k] = (T)sd[k];
        }
        for (int k = 0; k < BS_NSTATE; ++k) S[(size_t)k * sp.n + j] = (double)s[k];
        H[j] = spk ? 1.0 : 0.0;  // (out >= 0.5).astype(dftype)
        if (spk && sp.record_spikes) {
          net->out_gid.push_back((int32_t)(net->gid0[p] + j));
          net->out_step.push_back((int32_t)n);
        }
      }
    }
    if (vm_out)
      for (int k = 0; k < ntaps; ++k) {
        const int p = net->taps[k].first, j = net->taps[k].second;
        vm_out[t * ntaps + k] = net->st[p][j];
      }
    net->step++;
  }
  return BS_OK;
}

extern "C" int orc_run(orc_net* net, int64_t n_steps, double* vm_out) {
  if (!net || n_steps < 0) return fail(BS_EINVAL, "orc_run: bad arguments");
  net->out_gid.clear();
  net->out_step.clear();
  return net->dtype == BS_F64 ? run_t<double>(net, n_steps, vm_out) : run_t<float>(net, n_steps, vm_out);
}

extern "C" int orc_spike_count(orc_net* net, int64_t* n) {
  if (!net || !n) return fail(BS_EINVAL, "orc_spike_count: null argument");
  *n = (int64_t)net->out_gid.size();
  return BS_OK;
}

extern "C" int orc_copy_spikes(orc_net* net, int32_t* gid, int32_t* step, int64_t cap) {
  if (!net || !gid || !step) return fail(BS_EINVAL, "orc_copy_spikes: null argument");
  if (cap < (int64_t)net->out_gid.size()) return fail(BS_ECAPACITY, "orc_copy_spikes: cap too small");
  // already in (step, gid) order: steps ascend and populations/neurons are
  // visited in gid order within a step
  memcpy(gid, net->out_gid.data(), 4 * net->out_gid.size());
  memcpy(step, net->out_step.data(), 4 * net->out_step.size());
  return BS_OK;
}

extern "C" int orc_get_state(orc_net* net, int32_t pop, int32_t var, double* out) {
  if (!net || !out || pop < 0 || pop >= (int)net->pops.size() || var < -1 || var >= BS_NSTATE)
    return fail(BS_EINVAL, "orc_get_state: bad arguments");
  const int n = net->pops[pop].n;
  for (int i = 0; i < n; ++i) out[i] = var < 0 ? (double)net->ref[pop][i] : net->st[pop][(size_t)var * n + i];
  return BS_OK;
}

// Injected generator counts (the engine's bs_net_set_drive): channel c's
// generator steps step0 .. step0 + n_steps - 1 take counts[t][e] (rows of
// n_gen elements) instead of draws; n_steps = 0 removes them.
extern "C" int orc_set_drive(orc_net* net, int32_t c, int64_t step0, int64_t n_steps, int32_t n_gen,
                             const int32_t* counts) {
  if (!net || c < 0 || c >= (int)net->chans.size() || net->chans[c].kind != BS_CH_POISSON)
    return fail(BS_EINVAL, "orc_set_drive: channel is not a Poisson channel");
  if (step0 < 0 || n_steps < 0 || (n_steps > 0 && (!counts || n_gen <= 0)))
    return fail(BS_EINVAL, "orc_set_drive: bad arguments");
  int width = 0;
  for (int g : net->gen_index[c]) width = std::max(width, g + 1);
  if (n_steps > 0 && n_gen < width) return fail(BS_EINVAL, "orc_set_drive: rows narrower than the generator");
  net->drive[c].assign(counts, counts + (size_t)n_steps * (n_steps > 0 ? n_gen : 0));
  net->drive0[c] = step0;
  net->drive_n[c] = n_steps;
  net->drive_w[c] = n_gen;
  return BS_OK;
}

extern "C" int orc_get_weights(orc_net* net, int32_t proj, double* out) {
  if (!net || !out || proj < 0 || proj >= (int)net->projs.size())
    return fail(BS_EINVAL, "orc_get_weights: bad arguments");
  const OrcProj& P = net->projs[proj];
  for (size_t e = 0; e < P.w.size(); ++e) out[e] = P.w[e];
  return BS_OK;
}

// One model update of a single neuron with explicit inputs (the reference's
// unit tests drive neurons this way: add_delta_input + update()). in0 is the
// summed delta input (mV for iaf_psc_delta, pA for iaf_psc_exp/alpha, nS into
// g_ex for aeif_cond_alpha), in1 the aeif g_in input. state: BS_NSTATE values.
extern "C" int orc_neuron_step(int model, const double* params, double* state, int* ref,
                               double in0, double in1, int dtype) {
  std::vector<double> P(params, params + BS_NPARAM);
  if (dtype == BS_F32 && model != BS_AEIF_COND_ALPHA) {
    float s[BS_NSTATE];
    for (int k = 0; k < BS_NSTATE; ++k) s[k] = (float)state[k];
    bool spk = false;
    if (model == BS_IAF_PSC_DELTA) spk = upd_delta<float>(P, s[0], s[1], *ref, (float)in0, (float)P[34] + 0.0f);
    else if (model == BS_IAF_PSC_EXP) spk = upd_exp<float>(P, s, *ref, (float)in0, 0.0f);
    else if (model == BS_IAF_PSC_ALPHA) spk = upd_alpha<float>(P, s, *ref, (float)in0, 0.0f);
    else return -1;
    for (int k = 0; k < BS_NSTATE; ++k) state[k] = s[k];
    return spk ? 1 : 0;
  }
  bool spk = false;
  if (model == BS_IAF_PSC_DELTA) spk = upd_delta<double>(P, state[0], state[1], *ref, in0, P[34] + 0.0);
  else if (model == BS_IAF_PSC_EXP) spk = upd_exp<double>(P, state, *ref, in0, 0.0);
  else if (model == BS_IAF_PSC_ALPHA) spk = upd_alpha<double>(P, state, *ref, in0, 0.0);
  else if (model == BS_AEIF_COND_ALPHA) {
    bool unstable = false;
    spk = aeif::update(P, state, *ref, in0, in1, 0.0, &unstable);
    if (unstable) return -5;
  } else {
    return -1;
  }
  return spk ? 1 : 0;
}
