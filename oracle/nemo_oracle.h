/*
 * nemo_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a Boost-free C++ restatement of
 * NeMo's CPU backend (src/nemo/cpu/, src/nemo/ConnectivityMatrix.cpp,
 * src/nemo/StdpProcess.cpp, src/nemo/RNG.cpp, src/nemo/fixedpoint.cpp and the
 * six CPU neuron-model plugins).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the CPU
 * baseline.  The product library (nemo_amd) never links or calls it.
 *
 * All functions return 0 on success and a NEMO_* status code otherwise; the
 * message of the most recent failure is available from oracle_strerror().
 */
#ifndef NEMO_ORACLE_H
#define NEMO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_network* oracle_network_t;
typedef struct oracle_configuration* oracle_configuration_t;
typedef struct oracle_simulation* oracle_simulation_t;

const char* oracle_strerror(void);

/* --- network construction (src/nemo/NetworkImpl.cpp:27-128) --- */
oracle_network_t oracle_new_network(void);
void oracle_delete_network(oracle_network_t);
int oracle_add_neuron_type(oracle_network_t, const char* name, unsigned* type);
/* n neurons; args is n x nargs row-major (parameters then state variables) */
int oracle_add_neurons(oracle_network_t, unsigned type, size_t n,
		const unsigned* idx, unsigned nargs, const float* args);
int oracle_add_synapses(oracle_network_t, size_t n,
		const unsigned* source, const unsigned* target, const unsigned* delay,
		const float* weight, const unsigned char* plastic);
int oracle_neuron_count(oracle_network_t, unsigned* count);

/* --- configuration (src/nemo/ConfigurationImpl.cpp, StdpFunction.cpp) --- */
oracle_configuration_t oracle_new_configuration(void);
void oracle_delete_configuration(oracle_configuration_t);
int oracle_set_stdp_function(oracle_configuration_t,
		const float* prefire, size_t prefire_len,
		const float* postfire, size_t postfire_len,
		float minE, float maxE, float minI, float maxI);

/* --- simulation (src/nemo/cpu/Simulation.cpp) --- */
int oracle_new_simulation(oracle_network_t, oracle_configuration_t, oracle_simulation_t*);
void oracle_delete_simulation(oracle_simulation_t);
int oracle_step(oracle_simulation_t,
		const unsigned* fstim, size_t fstim_count,
		const unsigned* istim_nidx, const float* istim_current, size_t istim_count,
		const unsigned** fired, size_t* fired_count);
int oracle_apply_stdp(oracle_simulation_t, float reward);
int oracle_get_membrane_potential(oracle_simulation_t, unsigned neuron, float* v);
int oracle_get_neuron_state(oracle_simulation_t, unsigned neuron, unsigned var, float* val);
int oracle_get_neuron_parameter(oracle_simulation_t, unsigned neuron, unsigned param, float* val);
int oracle_set_neuron_state(oracle_simulation_t, unsigned neuron, unsigned var, float val);
int oracle_set_neuron_parameter(oracle_simulation_t, unsigned neuron, unsigned param, float val);
int oracle_get_synapses_from(oracle_simulation_t, unsigned source, const uint64_t** ids, size_t* len);
int oracle_get_synapse_target(oracle_simulation_t, uint64_t id, unsigned* target);
int oracle_get_synapse_delay(oracle_simulation_t, uint64_t id, unsigned* delay);
int oracle_get_synapse_weight(oracle_simulation_t, uint64_t id, float* weight);
int oracle_get_synapse_plastic(oracle_simulation_t, uint64_t id, unsigned char* plastic);
int oracle_elapsed_simulation(oracle_simulation_t, unsigned long* ms);
int oracle_reset_timer(oracle_simulation_t);

/* --- inspection hooks used by the parity tests --- */
/* number of neurons and the global index of each local (simulation) index */
int oracle_local_to_global(oracle_simulation_t, unsigned* gidx, size_t n);
/* float currents (E then I) delivered in the most recent step, in local order */
int oracle_get_currents(oracle_simulation_t, float* E, float* I, size_t n);
/* fixed-point weight of every synapse, in (source ascending, insertion) order */
int oracle_get_fixed_weights(oracle_simulation_t, int32_t* w, size_t n);
/* STDP accumulator of every synapse in the same order (0 for static synapses) */
int oracle_get_stdp_accumulators(oracle_simulation_t, int32_t* acc, size_t n);
/* 64-bit firing history in local order */
int oracle_get_recent_firing(oracle_simulation_t, uint64_t* hist, size_t n);
/* RNG state (4 words per neuron) in local order */
int oracle_get_rng_state(oracle_simulation_t, uint32_t* state, size_t n);
/* number of synaptic events delivered since construction */
uint64_t oracle_delivered_events(oracle_simulation_t);
/* OpenMP threads the oracle will use */
int oracle_num_threads(void);

/* state variable var of every neuron in local order (NaN if the model has fewer) */
int oracle_get_neuron_states(oracle_simulation_t, unsigned var, float* out, size_t n);

/* --- Brian's propagate (src/nemo/cpu/Simulation.cpp:115-137) --- */
/* the neurons with the given local indices fire; E and I point at the input
 * delivered now, in local order (valid until the next call) */
int oracle_propagate(oracle_simulation_t, const unsigned* fired, size_t nfired, const float** E, const float** I);

/* --- RNG primitives (src/nemo/RNG.cpp:9-63) --- */
void oracle_rng_seeds(unsigned first, unsigned last, uint32_t* out /* 4*(last-first+1) */);
uint32_t oracle_urand(uint32_t* state4);
float oracle_nrand(uint32_t* state4);
/* nrand of n generators (4 words each, advanced in place) */
void oracle_nrand_n(uint32_t* states, size_t n, float* out);
uint32_t oracle_mt19937_nth(uint32_t seed, unsigned n);

/* --- fixed point (src/nemo/fixedpoint.cpp:17-49) --- */
int oracle_fx_tofix(float f, unsigned fbits, int32_t* out);
float oracle_wfx_tofloat(int64_t v, unsigned fbits);

#ifdef __cplusplus
}
#endif

#endif
