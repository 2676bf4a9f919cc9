// bs_models.h -- one-step neuron updates executed per lane, restating the
// reference's update() bodies operation by operation (same association order,
// no contraction) so FP64 results are bit-identical to a sequential CPU
// restatement. Parameter-block layouts (double[40]) are filled by
// brainpy/state_amd/models.py with host-precomputed FP64 propagators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bs_math.h"

namespace bs {

// ---- parameter indices --------------------------------------------------
// iaf_psc_delta  (_nest_neuron/iaf_psc_delta.py:449-521)
enum { D_EL = 0, D_DECAY, D_CI, D_VTH, D_VRESET, D_VMIN, D_REFN, D_REFIN, D_RDEC0 };
// state: 0 V, 1 refractory_spike_buffer
// iaf_psc_exp    (_nest_neuron/iaf_psc_exp.py:552-682)
enum { X_EL = 0, X_THETA, X_VRESET_REL, X_P22, X_P21EX, X_P21IN, X_P20, X_P11EX, X_P11IN,
       X_IE, X_REFN, X_1MP11EX };
// state: 0 V, 1 i_syn_ex, 2 i_syn_in, 3 i_0, 4 i_1
// iaf_psc_alpha  (_nest_neuron/iaf_psc_alpha.py:558-684)
enum { A_EL = 0, A_THETA, A_VRESET_REL, A_P11EX, A_P21EX, A_P22EX, A_P11IN, A_P21IN, A_P22IN,
       A_EXPM1, A_P30, A_P31EX, A_P32EX, A_P31IN, A_P32IN, A_EPSC, A_IPSC, A_IE, A_REFN,
       A_VMIN_REL };
// state: 0 V, 1 I_syn_ex, 2 I_syn_in, 3 dI_syn_ex, 4 dI_syn_in, 5 y0
// aeif_cond_alpha (_nest_neuron/aeif_cond_alpha.py:554-747, _nest_base/utils.py:583-723)
enum { C_VPEAK = 0, C_VRESET, C_GL, C_CM, C_EEX, C_EIN, C_EL, C_DT_, C_TAUW, C_A, C_B, C_VTH,
       C_TAUEX, C_TAUIN, C_IE, C_ATOL, C_REFN, C_PSCEX, C_PSCIN, C_DT, C_MINH, C_VPEAKDET,
       C_DTSAFE, C_MAXIT, C_GLDT, C_NEGGL };
// state: 0 V, 1 dg_ex, 2 g_ex, 3 dg_in, 4 g_in, 5 w, 6 h, 7 I_stim

// Parameters are loaded once per launch into a per-model struct (registers),
// converted to T exactly as the per-step (T)P[i] casts would be.

// ---- iaf_psc_delta ------------------------------------------------------
// v = E_L + (V-E_L)*decay + (I/C)*tau*(1-decay)  [the last term is the
// constant D_CI without current inputs; with them I = (I_e + 0) + currents of
// this step, evaluated per step]; v += sum(delta); refractory holds V; spike
// iff v >= V_th.
template <typename T>
struct DeltaP {
  T EL, decay, cI, Vth, Vreset, Vmin, Cm, taum, one_m_decay;
  int refn, refin, has_vmin;
  const double* P;  // refractory decay table (refractory_input)
};
template <typename T>
__device__ __forceinline__ DeltaP<T> load_delta(const double* P) {
  DeltaP<T> p;
  p.EL = (T)P[D_EL];
  p.decay = (T)P[D_DECAY];
  p.cI = (T)P[D_CI];
  p.Vth = (T)P[D_VTH];
  p.Vreset = (T)P[D_VRESET];
  p.Vmin = (T)P[D_VMIN];
  p.Cm = (T)P[32];
  p.taum = (T)P[33];
  p.one_m_decay = (T)1 - p.decay;
  p.refn = (int)P[D_REFN];
  p.refin = P[D_REFIN] != 0.0;
  p.has_vmin = P[D_VMIN] > -1e300;
  p.P = P;
  return p;
}
// HAS_CUR: the population has current-input channels (a compile-time branch,
// so the common case carries no per-step division).
template <typename T, bool HAS_CUR>
__device__ __forceinline__ bool step_delta(const DeltaP<T>& p, T* s, int& r, T din, T i_total) {
  const T last_v = s[0];
  T cI = p.cI;
  if constexpr (HAS_CUR) cI = (i_total / p.Cm) * p.taum * p.one_m_decay;
  T v = p.EL + (last_v - p.EL) * p.decay + cI;
  v = v + din;
  if (p.refin) v = v + s[1];
  if (p.has_vmin) v = v > p.Vmin ? v : p.Vmin;
  const bool not_ref = (r == 0);
  const T v_post = not_ref ? v : last_v;
  if (p.refin) {
    // exp(-r dt / tau_m) for the current r, tabulated on the host.
    const T rdec = (T)p.P[D_RDEC0 + (r < 23 ? r : 23)];
    s[1] = not_ref ? (T)0 : (T)(s[1] + din * rdec);
  }
  int rr = not_ref ? r : r - 1;
  const bool spike = v_post >= p.Vth;
  r = spike ? p.refn : rr;
  s[0] = spike ? p.Vreset : v_post;
  return spike;
}

// ---- iaf_psc_exp (delta = 0, deterministic threshold) ---------------------
template <typename T>
struct ExpP {
  T EL, theta, Vreset_rel, P22, P21ex, P21in, P20, P11ex, P11in, Ie, one_m_P11ex;
  int refn;
};
template <typename T>
__device__ __forceinline__ ExpP<T> load_exp(const double* P) {
  ExpP<T> p;
  p.EL = (T)P[X_EL];
  p.theta = (T)P[X_THETA];
  p.Vreset_rel = (T)P[X_VRESET_REL];
  p.P22 = (T)P[X_P22];
  p.P21ex = (T)P[X_P21EX];
  p.P21in = (T)P[X_P21IN];
  p.P20 = (T)P[X_P20];
  p.P11ex = (T)P[X_P11EX];
  p.P11in = (T)P[X_P11IN];
  p.Ie = (T)P[X_IE];
  p.one_m_P11ex = (T)P[X_1MP11EX];
  p.refn = (int)P[X_REFN];
  return p;
}
template <typename T>
__device__ __forceinline__ bool step_exp(const ExpP<T>& p, T* s, int& r, T w_all, T cur) {
  T V_rel = s[0] - p.EL;
  T iex = s[1], iin = s[2];
  const T i0 = s[3], i1 = s[4];
  const bool not_ref = (r == 0);
  const T Vc = V_rel * p.P22 + iex * p.P21ex + iin * p.P21in + (p.Ie + i0) * p.P20;
  V_rel = not_ref ? Vc : V_rel;
  int rr = not_ref ? r : r - 1;
  iex = iex * p.P11ex;
  iin = iin * p.P11in;
  iex = iex + p.one_m_P11ex * i1;
  const T wex = w_all > (T)0 ? w_all : (T)0;
  const T win = w_all < (T)0 ? w_all : (T)0;
  iex = iex + wex;
  iin = iin + win;
  const bool spike = V_rel >= p.theta;
  r = spike ? p.refn : rr;
  V_rel = spike ? p.Vreset_rel : V_rel;
  s[0] = V_rel + p.EL;
  s[1] = iex;
  s[2] = iin;
  s[3] = cur;  // i_0 for the next step: x + sum_current_inputs (x = 0)
  s[4] = (T)0;
  return spike;
}

// ---- iaf_psc_alpha -------------------------------------------------------
template <typename T>
struct AlphaP {
  T EL, theta, Vreset_rel, P11ex, P21ex, P22ex, P11in, P21in, P22in, expm1, P30, P31ex, P32ex,
      P31in, P32in, epsc, ipsc, Ie, Vmin_rel;
  int refn, has_vmin;
};
template <typename T>
__device__ __forceinline__ AlphaP<T> load_alpha(const double* P) {
  AlphaP<T> p;
  p.EL = (T)P[A_EL];
  p.theta = (T)P[A_THETA];
  p.Vreset_rel = (T)P[A_VRESET_REL];
  p.P11ex = (T)P[A_P11EX];
  p.P21ex = (T)P[A_P21EX];
  p.P22ex = (T)P[A_P22EX];
  p.P11in = (T)P[A_P11IN];
  p.P21in = (T)P[A_P21IN];
  p.P22in = (T)P[A_P22IN];
  p.expm1 = (T)P[A_EXPM1];
  p.P30 = (T)P[A_P30];
  p.P31ex = (T)P[A_P31EX];
  p.P32ex = (T)P[A_P32EX];
  p.P31in = (T)P[A_P31IN];
  p.P32in = (T)P[A_P32IN];
  p.epsc = (T)P[A_EPSC];
  p.ipsc = (T)P[A_IPSC];
  p.Ie = (T)P[A_IE];
  p.Vmin_rel = (T)P[A_VMIN_REL];
  p.refn = (int)P[A_REFN];
  p.has_vmin = P[A_VMIN_REL] > -1e300;
  return p;
}
template <typename T>
__device__ __forceinline__ bool step_alpha(const AlphaP<T>& p, T* s, int& r, T w_all, T cur) {
  const T wex = w_all > (T)0 ? w_all : (T)0;
  const T win = w_all < (T)0 ? w_all : (T)0;
  T y3 = s[0] - p.EL;
  T Iex = s[1], Iin = s[2], dIex = s[3], dIin = s[4];
  const T y0 = s[5];
  const bool not_ref = (r == 0);
  T y3c = p.P30 * (y0 + p.Ie) + p.P31ex * dIex + p.P32ex * Iex + p.P31in * dIin + p.P32in * Iin +
          p.expm1 * y3 + y3;
  if (p.has_vmin) y3c = y3c > p.Vmin_rel ? y3c : p.Vmin_rel;
  y3 = not_ref ? y3c : y3;
  int rr = not_ref ? r : r - 1;
  Iex = p.P21ex * dIex + p.P22ex * Iex;
  dIex = dIex * p.P11ex;
  dIex = dIex + p.epsc * wex;
  Iin = p.P21in * dIin + p.P22in * Iin;
  dIin = dIin * p.P11in;
  dIin = dIin + p.ipsc * win;
  const bool spike = y3 >= p.theta;
  r = spike ? p.refn : rr;
  y3 = spike ? p.Vreset_rel : y3;
  s[0] = y3 + p.EL;
  s[1] = Iex;
  s[2] = Iin;
  s[3] = dIex;
  s[4] = dIin;
  s[5] = cur;  // y0 for the next step: x + sum_current_inputs (x = 0)
  return spike;
}

// ---- aeif_cond_alpha: adaptive RKF45 per lane ----------------------------
// Butcher tableau (Fehlberg): the coefficient values are the same double
// expressions the reference evaluates (utils.py:266-278); zero entries are
// skipped as _rk_weighted_sum does (utils.py:388-408).
struct AeifY {
  double V, dgex, gex, dgin, gin, w;
};

// Reciprocals of the divisors of the right-hand side (div_rc), once per step.
struct AeifR {
  double dts, cm, tex, tin, tw;
};

// Right-hand side; DIV(x, c, rc) is the division x / c (exact: plain or
// reciprocal-corrected, both give the IEEE quotient).
#define BS_AEIF_RHS(DIV)                                                                     \
  const double veff = refr ? P[C_VRESET] : ((y.V != y.V || y.V < P[C_VPEAK]) ? y.V : P[C_VPEAK]); \
  const double isex = y.gex * (veff - P[C_EEX]);                                            \
  const double isin = y.gin * (veff - P[C_EIN]);                                            \
  double ea = DIV(veff - P[C_VTH], P[C_DTSAFE], R.dts);                                     \
  ea = ea < -500.0 ? -500.0 : (ea > 500.0 ? 500.0 : ea);                                   \
  const double ispk = P[C_GLDT] * exp_p(ea);                                                \
  const double dvraw =                                                                      \
      DIV(P[C_NEGGL] * (veff - P[C_EL]) + ispk - isex - isin - y.w + P[C_IE] + istim, P[C_CM], R.cm); \
  k.V = refr ? 0.0 : dvraw;                                                                 \
  k.dgex = DIV(-y.dgex, P[C_TAUEX], R.tex);                                                 \
  k.gex = y.dgex - DIV(y.gex, P[C_TAUEX], R.tex);                                           \
  k.dgin = DIV(-y.dgin, P[C_TAUIN], R.tin);                                                 \
  k.gin = y.dgin - DIV(y.gin, P[C_TAUIN], R.tin);                                           \
  k.w = DIV(P[C_A] * (veff - P[C_EL]) - y.w, P[C_TAUW], R.tw);

__device__ __forceinline__ void aeif_rhs_ieee(const double* __restrict__ P, const AeifR& R, const AeifY& y,
                                              bool refr, double istim, AeifY& k) {
#define BS_DIV_IEEE(x, c, rc) ((x) / (c))
  BS_AEIF_RHS(BS_DIV_IEEE)
#undef BS_DIV_IEEE
}

// Every division of one evaluation through div_rc, straight-line; in the rare
// evaluation where one of them cannot use it, all are redone with x / c.
__device__ __forceinline__ void aeif_rhs(const double* __restrict__ P, const AeifR& R, const AeifY& y,
                                         bool refr, double istim, AeifY& k) {
  bool bad = false;
#define BS_DIV_RC(x, c, rc) div_rc((x), (c), (rc), bad)
  BS_AEIF_RHS(BS_DIV_RC)
#undef BS_DIV_RC
  if (bad) aeif_rhs_ieee(P, R, y, refr, istim, k);
}
#undef BS_AEIF_RHS

#define BS_AX(f, expr) t.f = y.f + h * (expr);
__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  return a > b ? a : b;
}

// The integration of one aeif_cond_alpha step in resumable pieces, so that
// the lanes of a wave can each be at a different step of the window: a lane
// whose step needs many substeps (a neuron crossing threshold needs hundreds)
// then holds up only its own later steps, not every lane's. Each piece is a
// slice of the reference's substep loop (utils.py:583-723) with the same
// operations in the same order; step_aeif below chains them into one step.
struct AeifLane {
  AeifY y, k0;  // state; k0 = f(y) of the coming substep
  double h, tl, istim;
  int it;
  bool spiked;
};

__device__ __forceinline__ AeifR aeif_recips(const double* __restrict__ P) {
  return AeifR{1.0 / P[C_DTSAFE], 1.0 / P[C_CM], 1.0 / P[C_TAUEX], 1.0 / P[C_TAUIN], 1.0 / P[C_TAUW]};
}

__device__ __forceinline__ void aeif_load(AeifLane& L, const double* s) {
  L.y = AeifY{s[0], s[1], s[2], s[3], s[4], s[5]};
  L.h = s[6];
  L.istim = s[7];
}

__device__ __forceinline__ void aeif_store(const AeifLane& L, double* s) {
  s[0] = L.y.V;
  s[1] = L.y.dgex;
  s[2] = L.y.gex;
  s[3] = L.y.dgin;
  s[4] = L.y.gin;
  s[5] = L.y.w;
  s[6] = L.h;
  s[7] = L.istim;
}

// Start of a step (the caller then evaluates k0 = f(y)).
__device__ __forceinline__ void aeif_begin(const double* __restrict__ P, AeifLane& L) {
  const double minh = P[C_MINH];
  L.h = L.h > minh ? L.h : minh;
  L.tl = 0.0;
  L.it = 0;
  L.spiked = false;
}

// Whether the step's substep loop runs another substep.
__device__ __forceinline__ bool aeif_more(const double* __restrict__ P, const AeifLane& L) {
  return L.it < (int)P[C_MAXIT] && L.tl < P[C_DT];
}

// One substep: the five further stages from k0, the embedded error, accept
// (state, instability guard, refractory clamp, spike/reset/adaptation) or
// reject, the new step size. Returns whether it was accepted; after an
// accepted substep k0 no longer matches y and the caller re-evaluates it
// before the next substep.
__device__ __forceinline__ bool aeif_substep(const double* __restrict__ P, const AeifR& R, AeifLane& L, int& r,
                                             int* unstable) {
  const double A10 = 1.0 / 4.0;
  const double A20 = 3.0 / 32.0, A21 = 9.0 / 32.0;
  const double A30 = 1932.0 / 2197.0, A31 = -7200.0 / 2197.0, A32 = 7296.0 / 2197.0;
  const double A40 = 439.0 / 216.0, A41 = -8.0, A42 = 3680.0 / 513.0, A43 = -845.0 / 4104.0;
  const double A50 = -8.0 / 27.0, A51 = 2.0, A52 = -3544.0 / 2565.0, A53 = 1859.0 / 4104.0,
               A54 = -11.0 / 40.0;
  const double B0 = 16.0 / 135.0, B2 = 6656.0 / 12825.0, B3 = 28561.0 / 56430.0, B4 = -9.0 / 50.0,
               B5 = 2.0 / 55.0;
  const double E0 = 16.0 / 135.0 - 25.0 / 216.0, E2 = 6656.0 / 12825.0 - 1408.0 / 2565.0,
               E3 = 28561.0 / 56430.0 - 2197.0 / 4104.0, E4 = -9.0 / 50.0 - -1.0 / 5.0,
               E5 = 2.0 / 55.0 - 0.0;
  const double dt = P[C_DT], minh = P[C_MINH], atol = P[C_ATOL];
  const int refn = (int)P[C_REFN];
  const double istim = L.istim;
  const AeifY& y = L.y;
  const AeifY& k0 = L.k0;
  double h = L.h;
  {
    const double rem = dt - L.tl;
    const double hm = h < rem ? h : rem;
    h = minh > hm ? minh : hm;
  }
  const bool refr = r > 0;
  AeifY k1, k2, k3, k4, k5, t;
#define BS_STAGE1(f) t.f = y.f + h * (A10 * k0.f);
  BS_STAGE1(V) BS_STAGE1(dgex) BS_STAGE1(gex) BS_STAGE1(dgin) BS_STAGE1(gin) BS_STAGE1(w)
  aeif_rhs(P, R, t, refr, istim, k1);
#define BS_STAGE2(f) t.f = y.f + h * (A20 * k0.f + A21 * k1.f);
  BS_STAGE2(V) BS_STAGE2(dgex) BS_STAGE2(gex) BS_STAGE2(dgin) BS_STAGE2(gin) BS_STAGE2(w)
  aeif_rhs(P, R, t, refr, istim, k2);
#define BS_STAGE3(f) t.f = y.f + h * (A30 * k0.f + A31 * k1.f + A32 * k2.f);
  BS_STAGE3(V) BS_STAGE3(dgex) BS_STAGE3(gex) BS_STAGE3(dgin) BS_STAGE3(gin) BS_STAGE3(w)
  aeif_rhs(P, R, t, refr, istim, k3);
#define BS_STAGE4(f) t.f = y.f + h * (A40 * k0.f + A41 * k1.f + A42 * k2.f + A43 * k3.f);
  BS_STAGE4(V) BS_STAGE4(dgex) BS_STAGE4(gex) BS_STAGE4(dgin) BS_STAGE4(gin) BS_STAGE4(w)
  aeif_rhs(P, R, t, refr, istim, k4);
#define BS_STAGE5(f) t.f = y.f + h * (A50 * k0.f + A51 * k1.f + A52 * k2.f + A53 * k3.f + A54 * k4.f);
  BS_STAGE5(V) BS_STAGE5(dgex) BS_STAGE5(gex) BS_STAGE5(dgin) BS_STAGE5(gin) BS_STAGE5(w)
  aeif_rhs(P, R, t, refr, istim, k5);
  AeifY yh, ye;
#define BS_HIGH(f) yh.f = y.f + h * (B0 * k0.f + B2 * k2.f + B3 * k3.f + B4 * k4.f + B5 * k5.f);
  BS_HIGH(V) BS_HIGH(dgex) BS_HIGH(gex) BS_HIGH(dgin) BS_HIGH(gin) BS_HIGH(w)
#define BS_ERR(f) ye.f = h * (E0 * k0.f + E2 * k2.f + E3 * k3.f + E4 * k4.f + E5 * k5.f);
  BS_ERR(V) BS_ERR(dgex) BS_ERR(gex) BS_ERR(dgin) BS_ERR(gin) BS_ERR(w)
  // scaled error |e| / (atol + 0 * max(|y0|, |y1|)), max over the six leaves
  double err;
  {
#define BS_SE(f) (__builtin_fabs(ye.f) / (atol + 0.0 * nanmax(__builtin_fabs(y.f), __builtin_fabs(yh.f))))
    err = BS_SE(V);
    err = nanmax(err, BS_SE(dgex));
    err = nanmax(err, BS_SE(gex));
    err = nanmax(err, BS_SE(dgin));
    err = nanmax(err, BS_SE(gin));
    err = nanmax(err, BS_SE(w));
  }
  if (err != err) err = __builtin_inf();
  const bool accept = (err <= 1.0) || (h <= minh);
  if (accept) {
    L.y = yh;
    L.tl = L.tl + h;
    // event_fn: instability guard, refractory clamp, spike/reset/adaptation
    if (L.y.V < -1e3 || L.y.w < -1e6 || L.y.w > 1e6) *unstable = 1;
    if (r > 0) L.y.V = P[C_VRESET];
    if (r <= 0 && L.y.V >= P[C_VPEAKDET]) {
      L.spiked = true;
      L.y.V = P[C_VRESET];
      L.y.w = L.y.w + P[C_B];
      if (refn > 0) r = refn + 1;
    }
  }
  const double errsg = err > 1e-30 ? err : 1e-30;
  const double inv = 1.0 / errsg;
  // one pow for both branches (err^(-1/5) after acceptance, err^(-1/4)
  // after rejection), so a wave with both kinds of lanes runs it once
  const double pw = pow_p(inv, accept ? 0.2 : 0.25);
  if (accept) {
    double fa;
    if (errsg <= 1e-30) {
      fa = 5.0;
    } else {
      fa = 0.9 * pw;
      fa = fa < 0.2 ? 0.2 : (fa > 5.0 ? 5.0 : fa);
    }
    const double hn = h * fa;
    L.h = minh > hn ? minh : hn;
  } else {
    double fr = 0.9 * pw;
    fr = fr < 0.2 ? 0.2 : (fr > 1.0 ? 1.0 : fr);
    const double hn = h * fr;
    L.h = minh > hn ? minh : hn;
  }
  ++L.it;
  return accept;
}

// End of a step: refractory count, the step's synaptic inputs into the
// conductance derivatives, the current input latched for the next step.
__device__ __forceinline__ void aeif_finish(const double* __restrict__ P, AeifLane& L, int& r, double wex_in,
                                            double win_in, double cur) {
  if (r > 0) r = r - 1;
  L.y.dgex = L.y.dgex + P[C_PSCEX] * wex_in;
  L.y.dgin = L.y.dgin + P[C_PSCIN] * win_in;
  L.istim = cur;  // I_stim for the next step: x + sum_current_inputs (x = 0)
}

// One simulation step of aeif_cond_alpha for one neuron. Returns spike flag;
// sets *unstable when the reference's guard would fire. k0 = f(y) is
// evaluated before the first substep and after each accepted one that does
// not end the step; a rejected substep leaves y and r unchanged, so its k0
// serves again (the values are those of evaluating k0 at the top of every
// substep).
__device__ __forceinline__ bool step_aeif(const double* __restrict__ P, double* s, int& r,
                                          double wex_in, double win_in, double cur, int* unstable) {
  const AeifR R = aeif_recips(P);
  AeifLane L;
  aeif_load(L, s);
  aeif_begin(P, L);
  aeif_rhs(P, R, L.y, r > 0, L.istim, L.k0);
  while (aeif_more(P, L)) {
    const bool acc = aeif_substep(P, R, L, r, unstable);
    if (acc && L.tl < P[C_DT]) aeif_rhs(P, R, L.y, r > 0, L.istim, L.k0);
  }
  aeif_finish(P, L, r, wex_in, win_in, cur);
  aeif_store(L, s);
  return L.spiked;
}

}  // namespace bs
