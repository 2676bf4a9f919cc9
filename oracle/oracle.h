/*
 * oracle.h -- CPU restatement of the reference's NEST-style Simulator step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (brainpy/state_amd) never calls it.
 *
 * It consumes the same bs_net_desc as libbrainspike_hip.so and advances the
 * network with the reference's own algorithm, step by step on one CPU thread:
 *   phase 1  every projection reads the previous step's holder vector, pushes
 *            it through a delay ring over the full pre-population vector
 *            (InputDelay, _nest_network/delay.py:114-128) or a per-edge delay
 *            (_kernel/delay_scatter.py semantics), and accumulates
 *            y[post] += w over its CSR rows in ascending pre order
 *            (_nest_network/projections.py:36-59, connectivity.py:39-54);
 *            generators contribute count * weight (event_proj.py:311-312);
 *   phase 2  every population sums its inputs in projection order
 *            (_base.py:312-383) and runs its model update; holders take the
 *            binarised spikes, generators their raw counts
 *            (simulator.py:1756-1788).
 */
#ifndef BS_ORACLE_H
#define BS_ORACLE_H
#include <stdint.h>
#include "../include/brainspike_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_net orc_net;

const char* orc_last_error(void);
int orc_create(const bs_net_desc* desc, orc_net** out);
int orc_destroy(orc_net* net);
int orc_run(orc_net* net, int64_t n_steps, double* vm_out);
int orc_spike_count(orc_net* net, int64_t* n);
int orc_copy_spikes(orc_net* net, int32_t* gid, int32_t* step, int64_t cap);
int orc_get_state(orc_net* net, int32_t pop, int32_t var, double* out);
int orc_get_weights(orc_net* net, int32_t proj, double* out);
int orc_set_drive(orc_net* net, int32_t c, int64_t step0, int64_t n_steps, int32_t n_gen, const int32_t* counts);

/* One update of a single neuron with explicit summed inputs; returns the
 * spike flag (0/1) or a negative error. */
int orc_neuron_step(int model, const double* params, double* state, int* ref, double in0,
                    double in1, int dtype);

/* One stdp_synapse edge update given this step's spikes and the traces
 * (which include this step's own spikes); returns the new weight. */
double orc_stdp_edge(double w, double pre_spike, double post_spike, double pre_trace,
                     double post_trace, const bs_stdp* S);

/* Elementwise primitives (the definitions the device must reproduce). */
void orc_threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t* o0,
                      uint32_t* o1);
uint32_t orc_poisson(uint32_t seed, uint32_t step, uint32_t elem, double lam);
double orc_exp(double x);
double orc_log(double x);
double orc_sin(double x);
double orc_pow(double x, double y);
/* connectivity samplers of the device path (definition in bs_conn.hip) */
int orc_conn_fixed_indegree(int64_t seed, int32_t n_pre, int32_t n_post, int32_t K, int32_t exclude_self,
                            int32_t* pre, int32_t* post);
int64_t orc_conn_bernoulli(int64_t seed, int32_t n_pre, int32_t n_post, double p, int32_t exclude_self,
                           int32_t* pre, int32_t* post, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
