/*
 * brainspike_hip.h -- C ABI of libbrainspike_hip.so, the MI355X (gfx950) engine
 * behind the NEST-style Simulator network step.
 *
 * Plain C: pointers, sizes, enums. No torch, HIP or C++ types cross this line.
 * Every entry point returns 0 on success or a negative BS_E* code; the message
 * of the last failure on the calling thread is returned by bs_last_error().
 * Nothing aborts the process. Calls on one bs_net handle must be serialised by
 * the caller; different handles are independent (each owns one HIP stream).
 *
 * Reference interfaces replaced (paths relative to the brainspike tree):
 *   bs_net_create   <- Simulator.create/connect building EventProjection,
 *                      EventPlasticProj, InputDelay, _SparseEventMatMul
 *                      (_nest_network/simulator.py:458-683, 1399-1514;
 *                       event_proj.py:77-206; event_plastic.py:276-340)
 *   bs_net_run      <- Simulator._run_window: brainstate for_loop over
 *                      Simulator.update phases 1 and 2 plus recorder taps
 *                      (_nest_network/simulator.py:1719-1788, 1841-1883)
 *   bs_net_reset    <- Simulator.reset_rollout (simulator.py:1803-1820)
 *   spikes/traces   <- SimulationResult.spikes / .trace (simulator.py:306-414)
 *   weights         <- EventPlasticProj.weight State, record_weight
 *                      (event_plastic.py:305, simulator.py:884-910)
 *   bs_conn_*       <- sample_fixed_indegree / sample_pairwise_bernoulli
 *                      (_nest_network/connectivity.py:81-140); with
 *                      bs_projection.on_device, csr_structure (:39-54)
 */
#ifndef BRAINSPIKE_HIP_H
#define BRAINSPIKE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BS_ABI_VERSION 9

/* ---- return codes ------------------------------------------------------ */
#define BS_OK 0
#define BS_EINVAL (-1)      /* invalid argument / descriptor             */
#define BS_ENOMEM (-2)      /* host or device allocation failed          */
#define BS_EHIP (-3)        /* HIP runtime error                         */
#define BS_ECAPACITY (-4)   /* a fixed-capacity buffer would overflow    */
#define BS_ENUMERIC (-5)    /* numerical instability flagged on device   */
#define BS_ENODEV (-6)      /* no usable gfx950 device                   */

/* ---- enums -------------------------------------------------------------- */
enum bs_dtype { BS_F64 = 0, BS_F32 = 1 };

enum bs_model {
  BS_IAF_PSC_DELTA = 1,   /* _nest_neuron/iaf_psc_delta.py:449-521 */
  BS_IAF_PSC_EXP = 2,     /* _nest_neuron/iaf_psc_exp.py:552-682   */
  BS_IAF_PSC_ALPHA = 3,   /* _nest_neuron/iaf_psc_alpha.py:558-684 */
  BS_AEIF_COND_ALPHA = 4  /* _nest_neuron/aeif_cond_alpha.py:627-747 */
};

/* Input channels of a population, summed in list order each step exactly as
 * Dynamics.sum_delta_inputs folds its dict (_base.py:312-383). */
enum bs_channel_kind {
  BS_CH_COUNT = 1,   /* homogeneous-weight events: value = S[k], S[k] = sequential
                        sum of k copies of the weight (CSR event matmul,
                        projections.py:36-59, with per-edge equal weights)       */
  BS_CH_FIXED = 2,   /* heterogeneous weights: exact 2^-F fixed-point sum       */
  BS_CH_POISSON = 3, /* one_to_one poisson_generator: value = count * weight
                        (event_proj.py:311-312, poisson_generator.py:501-576)    */
  BS_CH_CURRENT = 4, /* dc_generator current (pA) into the neuron's current-input
                        seam (simulator.py:1612-1745): value = weight (amplitude
                        times the connection weight) at steps t_min_step <= step
                        < t_max_step, else 0; summed in list order with the
                        other current channels, separately from the delta
                        inputs. iaf_psc_delta adds the sum to I_e in the same
                        step; iaf_psc_exp / iaf_psc_alpha / aeif_cond_alpha
                        latch it for the next step (i_0 / y0 / I_stim)          */
  BS_CH_AC = 5,      /* ac_generator current (ac_generator.py:264-325), a
                        current channel like BS_CH_CURRENT: value =
                        (ac[0] * sin(ac[1] * t + ac[2]) + ac[3]) * weight with
                        t = step * dt (ms), ac = {amplitude (pA), omega (rad/ms)
                        = 2 pi frequency / 1000, phase (rad), offset (pA)},
                        inside t_min_step <= step < t_max_step, else 0. The
                        sine is the engine's own (Cody-Waite reduction by
                        pi/2 and Taylor series), within a few ulp of the
                        exact sine for |argument| < 1.6e6 rad                   */
  BS_CH_SCHEDULE = 6 /* step_current_generator (step_current_generator.py:269-335),
                        a current channel like BS_CH_CURRENT for a whole
                        schedule: sched_len disjoint plateaus sorted by start,
                        value = sched_val[k] at steps sched[2k] <= step <
                        sched[2k + 1], else 0 (a binary search per window on
                        the device, any schedule length)                      */
};

/* Parameter block lengths (doubles) per model; layouts are documented in
 * brainpy/state_amd/models.py and mirrored by csrc/bs_models.h. */
#define BS_NPARAM 40
/* state variables per neuron, all stored in the network dtype except the
 * refractory counter (int32). Order per model in csrc/bs_models.h. */
#define BS_NSTATE 9

typedef struct bs_population {
  int32_t model;            /* enum bs_model                                    */
  int32_t n;                /* neurons                                          */
  double params[BS_NPARAM]; /* model parameters incl. host-precomputed FP64
                               propagators                                      */
  const double* state0;     /* [BS_NSTATE][n] initial state (row per variable);
                               NULL = model defaults from params                 */
  int32_t ch_begin;         /* this population's channels: channels[ch_begin ..  */
  int32_t ch_count;         /*                       ch_begin + ch_count)       */
  int32_t n_vm;             /* voltmeter taps: local indices vm_idx[0..n_vm)     */
  const int32_t* vm_idx;
  int32_t record_spikes;    /* 1: keep this population's spikes for the caller  */
  int32_t remote;           /* 1: the neurons live on another shard: no update,
                               no channels, no taps; their spikes arrive through
                               bs_net_inject_spikes and are delivered like local
                               ones                                             */
} bs_population;

typedef struct bs_channel {
  int32_t kind;            /* enum bs_channel_kind                               */
  int32_t receptor;        /* aeif_cond_alpha: 0 -> g_ex, 1 -> g_in; else 0       */
  double weight;           /* COUNT, POISSON, CURRENT, AC (1 = unscaled)         */
  int32_t max_count;       /* COUNT: largest event count per (post, step)        */
  int32_t fixed_shift;     /* FIXED: F, value = sum(q_e) * 2^-F                  */
  /* POISSON only: */
  double lam;              /* rate * dt / 1000                                   */
  int64_t t_min_step;      /* POISSON: active iff t_min_step < step <= t_max_step;
                              CURRENT: active iff t_min_step <= step < t_max_step */
  int64_t t_max_step;
  uint32_t seed;           /* Simulator._derive_seed(rng_seed, ordinal)          */
  int32_t delay_steps;     /* delay / dt                                         */
  const int32_t* gen_index;/* [n_post_pop] generator element feeding each post,
                              -1 if none; NULL = identity (POISSON, CURRENT, AC) */
  double ac[4];            /* AC: amplitude, omega (rad/ms), phase (rad), offset  */
  const int64_t* sched;    /* SCHEDULE: [2 * sched_len] plateau start, stop steps */
  const double* sched_val; /* SCHEDULE: [sched_len] plateau values (amplitude times
                              the connection weight)                             */
  int32_t sched_len;
  int32_t _pad_sched;
  double delay_frac;       /* POISSON: fraction of delay/dt in the network dtype
                              (InputDelay, delay.py:114-128): value =
                              (c[step-1-d] * (1 - f) + c[step-2-d] * f) * weight
                              in the network dtype; 0 = whole steps              */
} bs_channel;

/* STDP parameters (stdp_synapse.py:115-179): additive/multiplicative pair rule. */
typedef struct bs_stdp {
  double tau_plus, tau_minus, lambda_, alpha, mu_plus, mu_minus, Wmax;
  double decay_plus;   /* exp(-dt/tau_plus)  (event_plastic.py:411)           */
  double decay_minus;  /* exp(-dt/tau_minus)                                   */
} bs_stdp;

typedef struct bs_projection {
  int32_t pre_pop, post_pop; /* population indices                              */
  int32_t channel;           /* index into channels[] (global)                   */
  int32_t plastic;           /* 0 static, 1 stdp_synapse                          */
  int64_t n_edges;
  const int32_t* pre;        /* [n_edges] population-local pre index             */
  const int32_t* post;       /* [n_edges] population-local post index            */
  const int32_t* delay;      /* [n_edges] delay in steps, or NULL                 */
  int32_t delay_steps;       /* homogeneous delay in steps when delay == NULL    */
  int32_t on_device;         /* 1: pre, post, delay and weight are device
                                pointers (static projections only; e.g. from
                                bs_conn_*), read in place by bs_net_create      */
  const double* weight;      /* [n_edges] initial weights (FIXED/plastic) or NULL */
  bs_stdp stdp;
} bs_projection;

typedef struct bs_net_desc {
  int32_t abi_version;       /* BS_ABI_VERSION                                   */
  int32_t dtype;             /* enum bs_dtype                                    */
  double dt_ms;
  int32_t n_pops, n_channels, n_projs, max_window; /* max_window 0 = automatic   */
  const bs_population* pops;
  const bs_channel* channels;
  const bs_projection* projs;
  const double* tables;      /* concatenated S tables of all COUNT channels      */
  const int64_t* table_off;  /* [n_channels] offset into tables, -1 if none      */
  int64_t n_tables;
  int32_t tile_size;         /* post neurons per block: 32, 64, 128 or 256; 0 =
                                the largest giving >= 1024 tiles, and 32 when
                                even 64 gives fewer tiles than the GPU has CUs   */
  int32_t plastic_records;   /* plastic arrivals a block holds at once (0 =
                                what the LDS leaves, up to 16384); a window
                                with more runs in sub-windows. No counterpart
                                in the reference (a capacity, not a result)     */
} bs_net_desc;

typedef struct bs_net bs_net;

/* ---- library ------------------------------------------------------------ */
int bs_abi_version(void);
/* SHA-256 (hex) of the sources this library was compiled from, as the build
 * computed it (brainpy/state_amd/build.py source_hash); the Python loader
 * refuses a library whose hash differs from the sources beside it. No
 * counterpart in the reference, whose Python is its own source. */
const char* bs_source_hash(void);
const char* bs_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int bs_device_count(void);

/* ---- network ------------------------------------------------------------ */
/* Copies the descriptor (params, edges, delays, weights, tables) to device
 * memory the handle owns; builds CSR rows and the arrival ring. */
int bs_net_create(const bs_net_desc* desc, bs_net** out);
int bs_net_destroy(bs_net* net);
/* Restore initial state, clear rings, traces and plastic weights; step = 0. */
int bs_net_reset(bs_net* net);
/* Advance n_steps. vm_out: host [n_steps][total voltmeter taps] (f64) or NULL.
 * Spikes of populations with record_spikes are kept for bs_net_copy_spikes. */
int bs_net_run(bs_net* net, int64_t n_steps, double* vm_out);
/* Steps taken since create / reset. */
int bs_net_step(bs_net* net, int64_t* step);
/* Spikes recorded by the last bs_net_run, sorted by (step, gid). */
int bs_net_spike_count(bs_net* net, int64_t* n);
int bs_net_copy_spikes(bs_net* net, int32_t* gid, int32_t* step, int64_t cap);
/* Neuron state after the last step: var in [0, BS_NSTATE) (values as f64) or
 * -1 (refractory step counter). Replaces reading a model's state
 * (iaf_psc_delta.V etc., _base.py) and SimulationResult.trace. */
int bs_net_get_state(bs_net* net, int32_t pop, int32_t var, double* out);
/* Overwrite one state variable of a population (var as in bs_net_get_state;
 * values rounded to the network dtype, the counter to int32): the reference's
 * state assignment before a run (model.V.value = ..., Simulator.cont). */
int bs_net_set_state(bs_net* net, int32_t pop, int32_t var, const double* in);

/* ---- whole-engine checkpoint / resume ------------------------------------
 * Replaces the carried state of Simulator.cont (simulator.py:1822-1839) and a
 * BPTT restart from (state0, n0) (bptt.py:536-571): a snapshot of everything
 * the next step depends on -- neuron state and refractory counters, the
 * spike buckets of the delay horizon, the arrival rings, spike history and
 * plastic traces, weights and last-arrival steps, and the step counter (the position
 * of every counter-based random stream). Restoring it into a handle created
 * from the same descriptor continues the run bit for bit. The snapshot is an
 * opaque host byte string of bs_net_state_bytes() bytes; a snapshot of a
 * different network is rejected (BS_EINVAL). */
int bs_net_state_bytes(bs_net* net, int64_t* bytes);
int bs_net_save_state(bs_net* net, void* buf, int64_t cap);
int bs_net_load_state(bs_net* net, const void* buf, int64_t size);

/* ---- device-resident, stream-ordered run -----------------------------------
 * What an XLA FFI handler (or torch op) needs from Simulator._run_window
 * (simulator.py:1841-1883): the run is queued on the caller's HIP stream
 * `stream` (a hipStream_t; NULL = the null stream) and writes only device
 * buffers the caller owns; the call returns once the work is queued.
 *   raster  uint32 [n_steps][ceil(Ntot / 32)] (device): bit g % 32 of word
 *           g / 32 of row t is set iff neuron gid g fired at step t of the
 *           run, every population included (the reference's (T, n) 0/1
 *           raster, packed); the call clears it first
 *   vm      f64 [n_steps][n_taps] (device) or NULL: the voltmeter taps in
 *           caller order (bs_net_run's vm_out, on the device)
 *   status  int32 [1] (device) or NULL: 0, or the bs_net_run error class of
 *           a device-side failure (BS_ECAPACITY, BS_ENUMERIC)
 * The handle's own stream joins the caller's stream before and after the
 * run (events), so calls on one handle stay ordered with each other. */
int bs_net_run_device(bs_net* net, int64_t n_steps, void* stream, uint32_t* raster, double* vm,
                      int32_t* status);
/* Drive injection for parity (the contract's bs_net_set_drive): Poisson
 * channel `channel` takes its generator's counts for the generator steps
 * step0 .. step0 + n_steps - 1 from counts[t][e] (n_steps rows of n_gen >=
 * the generator's size, values in [0, 65535]) instead of drawing them; the
 * generator's activity window still applies (poisson_generator.py:570-576:
 * where(t_min < step <= t_max, counts, 0)), and a count drawn at step n
 * arrives at n + 1 + delay as before. Replaces the reference's
 * jax.random.poisson draws (poisson_generator.py:501-569), so its own counts
 * can be fed in for a run against its raster. n_steps = 0 removes the
 * injection. Steps outside the range draw from the channel's stream. */
int bs_net_set_drive(bs_net* net, int32_t channel, int64_t step0, int64_t n_steps, int32_t n_gen,
                     const int32_t* counts);
/* Plastic weights of projection proj in the reference's CSR (sorted-by-pre,
 * stable) edge order (event_plastic.py:293-296). */
int bs_net_get_weights(bs_net* net, int32_t proj, double* out);
/* record_weight taps (simulator.py:884-910): every following bs_net_run
 * copies the weights of the listed plastic projections after every step (in
 * the order of bs_net_get_weights) on the device, into host memory once per
 * chunk; n = 0 removes the taps. bs_net_copy_weight_trace(k) returns tap k's
 * trajectory of the last run, [n_steps][n_edges] f64. */
int bs_net_set_weight_taps(bs_net* net, int32_t n, const int32_t* projs);
int bs_net_copy_weight_trace(bs_net* net, int32_t k, double* out);
/* Per-kernel HIP-event timing of the following runs (off by default: the
 * events add launch overhead). */
int bs_net_set_timing(bs_net* net, int on);
/* Timing of the last run: device milliseconds spent in the delivery kernels,
 * the update kernels and the whole run, plus kernel launches per kind. */
int bs_net_last_timing(bs_net* net, double* deliver_ms, double* update_ms,
                       double* total_ms, int64_t* n_deliver, int64_t* n_update);
/* Algorithmic bytes moved by the last run (roofline numerators). */
int bs_net_last_bytes(bs_net* net, double* deliver_bytes, double* update_bytes,
                      int64_t* n_events);
/* Steps per window (one k_window launch) of this handle: a spike emitted in
 * one window reaches no step of the same window. */
int bs_net_window(bs_net* net, int32_t* W);

/* ---- shards: the spike-list exchange between windows ----------------------
 *
 * Replaces the min-delay communication of a distributed NEST run, which the
 * reference does not have (its Simulator is one process); the per-window
 * spike list is the exchange format: after every window each shard hands the
 * (gid, step) list of its own populations' spikes to the others, which inject
 * the spikes of their remote populations before their next window. The
 * shards' windows must have the same length (bs_net_window) and advance in
 * lockstep. gid in the global numbering (the populations' order in the
 * descriptor, remote ones included); every step in the window this handle
 * ran last; all of a window's spikes in one call, sorted by (step, gid) as
 * bs_net_copy_spikes returns them. */
int bs_net_inject_spikes(bs_net* net, int64_t n, const int32_t* gid, const int32_t* step);

/* ---- connectivity sampled on the device ------------------------------------
 *
 * Replaces the random rule samplers sample_fixed_indegree (with multapses) and
 * sample_pairwise_bernoulli (_nest_network/connectivity.py:81-140), which the
 * reference runs as a host loop of jax.random calls. Each edge is a pure
 * function of the seed and its position: key = (seed & 0xFFFFFFFF,
 * 0xE36F6E6E ^ (seed >> 32)); Threefry-2x32-20.
 *   fixed in-degree: post j, slot k < K draws (x0, x1) = threefry(key, (j, k)),
 *     c = floor((x1 * 2^32 + x0) * n_cand / 2^64); with exclude_self
 *     n_cand = n_pre - 1 and pre = c + (c >= j), else n_cand = n_pre, pre = c.
 *     Edge e = j * K + k (post-major, the reference's order).
 *   pairwise Bernoulli: pair (i, j) is connected iff x0 of threefry(key, (i, j))
 *     is below ceil(p * 2^32) (the uniform x0 * 2^-32 < p), i == j skipped with
 *     exclude_self; edges in row-major (pre, post) order.
 * pre_map / post_map (or NULL) translate the sampled segment-local indices to
 * population-local ones. All arrays are DEVICE pointers on the current device;
 * the calls return after the device work is done. */
int bs_conn_fixed_indegree(int64_t seed, int32_t n_pre, int32_t n_post, int32_t K, int32_t exclude_self,
                           const int32_t* pre_map, const int32_t* post_map, int32_t* pre, int32_t* post);
/* Count pass: row_ptr (device, n_pre + 1 entries) receives the exclusive prefix
 * of the per-row edge counts; n_edges and max_indegree (largest number of
 * edges into one post) are host outputs. */
int bs_conn_bernoulli_count(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                            int64_t* row_ptr, int64_t* n_edges, int32_t* max_indegree);
/* The count pass without any allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or allocation-free serving loops). Same arguments and
 * results as bs_conn_bernoulli_count, bit for bit; its partial results (row
 * and post counts, chunk sums) live in a fixed 16 MiB device array of the
 * library that every GPU holds until the process ends, so n_pre and n_post
 * are limited to 2^20 each (BS_EINVAL beyond). Every call on one GPU shares
 * that array: calls that overlap on one GPU race. The caller still provides
 * row_ptr (device, n_pre + 1 entries). */
int bs_conn_bernoulli_count_noalloc(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                                    int64_t* row_ptr, int64_t* n_edges, int32_t* max_indegree);
/* Fill pass: the edges of the count pass with the same arguments, written
 * from row_ptr into pre / post (n_edges entries each). */
int bs_conn_bernoulli_fill(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                           const int64_t* row_ptr, const int32_t* pre_map, const int32_t* post_map,
                           int32_t* pre, int32_t* post);

/* ---- surrogate-gradient BPTT of a recurrent iaf_psc_exp network ----------
 *
 * Replaces the kernel learner's episode gradient: BPTT._value_and_grad over
 * checkpointed_scan(Program.step) (_kernel/learn/bptt.py:437-491) for an
 * iaf_psc_exp population (_kernel/models/iaf_psc_exp.py:120-151) with
 * SpikeRule.check / count_ste (_kernel/spiking.py:65-79, 364-488),
 * detach_reset = True, a fast-sigmoid surrogate on the margin
 * (V - V_th) / (V_th - V_reset), and the rate_mse / membrane_mse readouts
 * (bptt.py:170-215). Neurons 0..n_exc-1 are excitatory (their spikes and the
 * input sources feed I_syn_ex), n_exc..N-1 inhibitory (they feed I_syn_in).
 * An event emitted at step n reaches its target at step n + delay.
 * Weights are dense row-major float32 matrices (pre, post) in pA; absent
 * edges hold 0. Array arguments of forward/backward are DEVICE pointers; the
 * calls return after the work has finished. */
#define BS_LOSS_RATE_MSE 0
#define BS_LOSS_MEMBRANE_MSE 1

typedef struct bs_bptt bs_bptt;

typedef struct bs_bptt_desc {
  int32_t abi_version;
  int32_t n_in, n_exc, n_inh;   /* input sources, excitatory, inhibitory     */
  int32_t batch, n_steps;
  int32_t delay_steps;          /* recurrent delay, >= 1                     */
  int32_t delay_in_steps;       /* input delay, >= 1                         */
  int32_t checkpoint;           /* steps between stored states, >= 1         */
  int32_t loss;                 /* BS_LOSS_*                                 */
  int32_t ref_steps;            /* refractory steps: the counter after a spike */
  int32_t gemm;                 /* transposed delivery: 0 = bf16 three-term
                                   split (hi*hi + hi*lo + lo*hi, fp32
                                   accumulate), 1 = fp32                      */
  double dt_ms;
  double E_L, V_th, V_reset, I_e;                 /* mV, mV, mV, pA          */
  double P22, P21ex, P21in, P11ex, P11in, P20;    /* exact propagators       */
  double surrogate_alpha;       /* slope of the fast sigmoid on the margin   */
  int32_t replay;               /* reverse pass, per checkpoint segment: -1
                                   auto (1 when the summed inputs take at most
                                   32 GiB), 0 recompute the segment with its
                                   event gathers, 1 the forward pass stores
                                   every step's summed inputs and the segment
                                   is replayed from them (same bits)          */
  int32_t _pad;
} bs_bptt_desc;

int bs_bptt_create(const bs_bptt_desc* desc, bs_bptt** out);
int bs_bptt_destroy(bs_bptt* h);
/* w_in [n_in, N], w_rec [N, N], in_counts uint8 [n_steps, batch, n_in]
 * (events of each source per step), v0 [batch, N] initial V_m or null (E_L),
 * target [N] (Hz for rate_mse, mV for membrane_mse). Writes the loss. */
int bs_bptt_forward(bs_bptt* h, const float* w_in, const float* w_rec, const uint8_t* in_counts,
                    const float* v0, const float* target, double* loss);
/* Gradients of the last forward's loss: g_in [n_in, N], g_rec [N, N]. */
int bs_bptt_backward(bs_bptt* h, float* g_in, float* g_rec);
/* Host copies from the last forward: spike bitmaps uint64 [n_steps, batch,
 * ceil(N/64)] (bit j%64 of word j/64) and spike counts float [batch, N];
 * either pointer may be null. */
int bs_bptt_spikes(bs_bptt* h, uint64_t* bitmaps, float* counts);

/* ---- test hooks: device arithmetic evaluated elementwise ------------------
 * No counterpart in the reference; they let tests compare the device's
 * exp/log/pow, Threefry-2x32 and Poisson sampler bit-for-bit with the C
 * oracle. op: 0 exp, 1 log, 2 pow(a, b), 3 a / b through the reciprocal-
 * corrected division the aeif_cond_alpha update uses (IEEE quotient bits). */
int bs_debug_math(int32_t op, int64_t n, const double* a, const double* b, double* out);
int bs_debug_threefry(int64_t n, const uint32_t* k0, const uint32_t* k1, const uint32_t* c0,
                      const uint32_t* c1, uint32_t* o0, uint32_t* o1);
int bs_debug_poisson(int64_t n, const uint32_t* seed, const uint32_t* step, const uint32_t* elem,
                     double lam, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* BRAINSPIKE_HIP_H */
