/*
 * nemo_hip.h -- C ABI of libnemo_hip.so, the MI355X (gfx950) HIP simulation
 * backend for the NeMo spiking-network simulator.
 *
 * The functions mirror NeMo's C API (src/nemo.h, implemented in
 * src/nemo/nemo_c.cpp) with the same names, argument meaning and status
 * codes, so a C or FFI caller of NeMo links against this library instead.
 * The backend factory entry points at the end mirror the dlopen'ed CUDA
 * backend ABI (src/nemo/cuda/create_simulation.hpp:14-22,
 * src/nemo/cuda/devices.hpp:21-56).  Citations are relative to the NeMo
 * v0.7.2 source tree.
 *
 * Error handling: every nemo_status_t function returns NEMO_OK (0) on
 * success; otherwise nemo_strerror() describes the most recent error
 * (src/nemo/nemo_c.cpp:20-60).  Handles are not thread-safe; one host thread
 * drives one simulation (as in the reference).
 *
 * Entry points marked "extension" have no counterpart in NeMo's C API; each
 * says which reference behaviour it generalises.
 */
#ifndef NEMO_HIP_H
#define NEMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: src/nemo/types.h:31-61 */
typedef int nemo_status_t;
#define NEMO_OK 0
#define NEMO_CUDA_INVOCATION_ERROR 1
#define NEMO_CUDA_ASSERTION_FAILURE 2
#define NEMO_CUDA_MEMORY_ERROR 3
#define NEMO_CUDA_ERROR 4
#define NEMO_API_UNSUPPORTED 5
#define NEMO_INVALID_INPUT 6
#define NEMO_BUFFER_OVERFLOW 7
#define NEMO_BUFFER_UNDERFLOW 8
#define NEMO_LOGIC_ERROR 9
#define NEMO_ALLOCATION_ERROR 10
#define NEMO_MPI_ERROR 11
#define NEMO_DL_ERROR 12
#define NEMO_IO_ERROR 13
#define NEMO_UNKNOWN_ERROR 14

/* Backends: src/nemo/types.h:63-66, plus the HIP backend this library provides */
typedef unsigned backend_t;
#define NEMO_BACKEND_CUDA 0
#define NEMO_BACKEND_CPU 1
#define NEMO_BACKEND_HIP 2

typedef uint64_t synapse_id; /* src/nemo/types.h:70: (source << 32) | index */

typedef struct nemo_network_class* nemo_network_t;
typedef struct nemo_configuration_class* nemo_configuration_t;
typedef struct nemo_simulation_class* nemo_simulation_t;

/* ---------------------------------------------------------------- library */

/* src/nemo.h:56-57 */
const char* nemo_version(void);
/* src/nemo.h:770-790 (nemo_c.cpp:20-60) */
const char* nemo_strerror(void);
/* src/nemo.h:72-74: directory searched for <Model>.ini and hip/lib<Model>.so */
nemo_status_t nemo_add_plugin_path(const char* dir);
/* src/nemo.h:88-95 (nemo_cuda_device_count / _description), for HIP devices */
nemo_status_t nemo_hip_device_count(unsigned* count);
nemo_status_t nemo_hip_device_description(unsigned device, const char** descr);

/* ---------------------------------------------------------- configuration */

/* src/nemo.h:105-110 */
nemo_configuration_t nemo_new_configuration(void);
void nemo_delete_configuration(nemo_configuration_t);
/* src/nemo.h:114-115 */
nemo_status_t nemo_log_stdout(nemo_configuration_t);
/* src/nemo.h:148-157 */
nemo_status_t nemo_set_stdp_function(nemo_configuration_t,
		float prefire_fn[], size_t prefire_len,
		float postfire_fn[], size_t postfire_len,
		float min_excitatory_weight, float max_excitatory_weight,
		float min_inhibitory_weight, float max_inhibitory_weight);
/* src/nemo.h:160-162.  Accepted, but this library only builds HIP simulations:
 * nemo_new_simulation then fails with NEMO_API_UNSUPPORTED. */
nemo_status_t nemo_set_cpu_backend(nemo_configuration_t);
/* src/nemo.h:166-168 (nemo_set_cuda_backend), for the HIP backend; dev -1 = default */
nemo_status_t nemo_set_hip_backend(nemo_configuration_t, int dev);
/* src/nemo.h:171-173 (nemo_cuda_device), for the HIP backend */
nemo_status_t nemo_hip_device(nemo_configuration_t, int* dev);
/* src/nemo.h:177-185 */
nemo_status_t nemo_backend(nemo_configuration_t, backend_t* backend);
nemo_status_t nemo_backend_description(nemo_configuration_t, const char** descr);
/* src/nemo.h:189-191 */
nemo_status_t nemo_set_write_only_synapses(nemo_configuration_t);

/* ------------------------------------------------------------ construction */

/* src/nemo.h:215-221 */
nemo_network_t nemo_new_network(void);
void nemo_delete_network(nemo_network_t);
/* src/nemo.h:235-239 */
nemo_status_t nemo_add_neuron_type(nemo_network_t, const char* name, unsigned* neuron_type);
/* src/nemo.h:245-251 (deprecated in the reference; kept for drop-in) */
nemo_status_t nemo_add_neuron_iz(nemo_network_t, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);
/* src/nemo.h:265-268: args = parameters followed by state variables */
nemo_status_t nemo_add_neuron(nemo_network_t, unsigned type, unsigned idx,
		unsigned nargs, float args[]);
/* src/nemo.h:278-286 */
nemo_status_t nemo_add_synapse(nemo_network_t, unsigned source, unsigned target,
		unsigned delay, float weight, unsigned char is_plastic, synapse_id* id);
/* src/nemo.h:290-292 */
nemo_status_t nemo_neuron_count(nemo_network_t, unsigned* ncount);

/* Extension: n neurons at once, args is n x nargs row-major.  Same effect as
 * n calls of nemo_add_neuron (the Python binding's vector form,
 * src/api/python/nemo_py.cpp add_neuron overloads). */
nemo_status_t nemo_add_neurons(nemo_network_t, unsigned type, size_t n,
		const unsigned idx[], unsigned nargs, const float args[]);
/* Extension: n synapses at once; plastic and ids may be NULL.  Same effect as
 * n calls of nemo_add_synapse (src/api/python/nemo_py.cpp add_synapse vector form). */
nemo_status_t nemo_add_synapses(nemo_network_t, size_t n,
		const unsigned source[], const unsigned target[], const unsigned delay[],
		const float weight[], const unsigned char plastic[], synapse_id ids[]);

/* src/nemo.h:399-420 (network getters) and 613-680 (network setters) */
nemo_status_t nemo_get_neuron_state_n(nemo_network_t, unsigned neuron, unsigned var, float* val);
nemo_status_t nemo_get_neuron_parameter_n(nemo_network_t, unsigned neuron, unsigned param, float* val);
nemo_status_t nemo_set_neuron_n(nemo_network_t, unsigned idx, unsigned nargs, float args[]);
nemo_status_t nemo_set_neuron_state_n(nemo_network_t, unsigned neuron, unsigned var, float val);
/* src/nemo.h:640-643 (deprecated in the reference; kept for drop-in):
 * Izhikevich neuron, a, b, c, d, u, v, sigma */
nemo_status_t nemo_set_neuron_iz_n(nemo_network_t, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);
nemo_status_t nemo_set_neuron_parameter_n(nemo_network_t, unsigned neuron, unsigned param, float val);
/* src/nemo.h:467-567 (network synapse getters) */
nemo_status_t nemo_get_synapse_source_n(nemo_network_t, synapse_id, unsigned* source);
nemo_status_t nemo_get_synapse_target_n(nemo_network_t, synapse_id, unsigned* target);
nemo_status_t nemo_get_synapse_delay_n(nemo_network_t, synapse_id, unsigned* delay);
nemo_status_t nemo_get_synapse_weight_n(nemo_network_t, synapse_id, float* weight);
nemo_status_t nemo_get_synapse_plastic_n(nemo_network_t, synapse_id, unsigned char* plastic);
nemo_status_t nemo_get_synapses_from_n(nemo_network_t, unsigned source, synapse_id* synapses[], size_t* len);

/* -------------------------------------------------------------- simulation */

/* src/nemo.h:311-317; the simulation is built by the backend the
 * configuration selects (src/nemo/nemo.cpp:109-147) */
nemo_simulation_t nemo_new_simulation(nemo_network_t, nemo_configuration_t);
void nemo_delete_simulation(nemo_simulation_t);
/* src/nemo.h:347-352.  fired is borrowed until the next step. */
nemo_status_t nemo_step(nemo_simulation_t,
		unsigned fstim_nidx[], size_t fstim_count,
		unsigned istim_nidx[], float istim_current[], size_t istim_count,
		unsigned* fired[], size_t* fired_count);
/* src/nemo.h:356-358 */
nemo_status_t nemo_apply_stdp(nemo_simulation_t, float reward);
/* src/nemo.h:380-382 and 437-458 */
nemo_status_t nemo_get_membrane_potential(nemo_simulation_t, unsigned neuron, float* v);
nemo_status_t nemo_get_neuron_state_s(nemo_simulation_t, unsigned neuron, unsigned var, float* val);
nemo_status_t nemo_get_neuron_parameter_s(nemo_simulation_t, unsigned neuron, unsigned param, float* val);
/* src/nemo.h:625-732 (simulation setters) */
nemo_status_t nemo_set_neuron_s(nemo_simulation_t, unsigned idx, unsigned nargs, float args[]);
nemo_status_t nemo_set_neuron_state_s(nemo_simulation_t, unsigned neuron, unsigned var, float val);
/* Extension: vector forms of nemo_set_neuron_s (args n x nargs row-major),
 * nemo_get/set_neuron_state_s and nemo_get/set_neuron_parameter_s over n
 * neurons, as the Python binding's list arguments (src/api/python/
 * nemo_py.cpp); same errors as n single calls, raised before anything is
 * changed.  For a neuron listed twice the last value wins. */
nemo_status_t nemo_set_neurons_s(nemo_simulation_t, size_t n, const unsigned idx[], unsigned nargs,
		const float args[]);
nemo_status_t nemo_get_neuron_states_s(nemo_simulation_t, size_t n, const unsigned idx[], unsigned var, float out[]);
nemo_status_t nemo_get_neuron_parameters_s(nemo_simulation_t, size_t n, const unsigned idx[], unsigned param,
		float out[]);
nemo_status_t nemo_set_neuron_states_s(nemo_simulation_t, size_t n, const unsigned idx[], unsigned var,
		const float val[]);
nemo_status_t nemo_set_neuron_parameters_s(nemo_simulation_t, size_t n, const unsigned idx[], unsigned param,
		const float val[]);
/* src/nemo.h:655-658 (deprecated in the reference; kept for drop-in) */
nemo_status_t nemo_set_neuron_iz_s(nemo_simulation_t, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);
nemo_status_t nemo_set_neuron_parameter_s(nemo_simulation_t, unsigned neuron, unsigned param, float val);
/* src/nemo.h:478-581 (simulation synapse getters) */
nemo_status_t nemo_get_synapse_source_s(nemo_simulation_t, synapse_id, unsigned* source);
nemo_status_t nemo_get_synapse_target_s(nemo_simulation_t, synapse_id, unsigned* target);
nemo_status_t nemo_get_synapse_delay_s(nemo_simulation_t, synapse_id, unsigned* delay);
nemo_status_t nemo_get_synapse_weight_s(nemo_simulation_t, synapse_id, float* weight);
nemo_status_t nemo_get_synapse_plastic_s(nemo_simulation_t, synapse_id, unsigned char* plastic);
nemo_status_t nemo_get_synapses_from_s(nemo_simulation_t, unsigned source, synapse_id* synapses[], size_t* len);
/* src/nemo.h:753-763 */
nemo_status_t nemo_elapsed_wallclock(nemo_simulation_t, unsigned long* ms);
nemo_status_t nemo_elapsed_simulation(nemo_simulation_t, unsigned long* ms);
nemo_status_t nemo_reset_timer(nemo_simulation_t);

/* Extension: run nsteps steps without stimulus and without a host round trip
 * per step (steps are captured into HIP graphs).  Equivalent to nsteps calls
 * of nemo_step with no stimulus (src/nemo/SimulationBackend.cpp:22-30) plus
 * nemo_apply_stdp(reward) after every stdp_period steps when stdp_period > 0
 * (the reference benchmark loop, src/examples/common.cpp:72-90).  The firing
 * of every step is returned: counts[s] neurons fired in step s, and their
 * indices, each step in ascending simulation order, are concatenated in
 * *fired (borrowed until the next call).  counts may be NULL to skip the
 * raster readout; *events receives the synaptic events delivered.  For a
 * batch (nemo_hip_new_batch) counts holds one entry per step and network. */
nemo_status_t nemo_hip_run(nemo_simulation_t, unsigned nsteps,
		unsigned stdp_period, float reward,
		unsigned counts[], unsigned* fired[], size_t* fired_len, uint64_t* events);
/* Extension (no counterpart in the reference): nemo_hip_run with the raster
 * read back and the fired neurons written straight into the caller's
 * fired_buf (fired_cap entries; *fired_len of them used), as they are decoded
 * beside the later steps, so that no copy follows the run.  Neurons beyond
 * fired_cap continue in the simulation's own list, *spill (*spill_len
 * entries), valid until the next call on the simulation. */
nemo_status_t nemo_hip_run_into(nemo_simulation_t, unsigned nsteps, unsigned stdp_period, float reward,
		unsigned counts[], unsigned fired_buf[], size_t fired_cap, size_t* fired_len, unsigned* spill[],
		size_t* spill_len, uint64_t* events);

/* Extension: a batch of independent networks (trials, parameter sweeps)
 * advanced together in the same kernel launches.  Each network keeps its own
 * neuron indices, its own RNG seeds (as in a simulation of it alone) and its
 * own firing output; neuron types are shared by name, and the configuration
 * (STDP function, fixed-point format, device) is common.  The handle is a
 * simulation: nemo_hip_run advances it (counts[s*count + k] neurons of network
 * k fired in step s; within a step the fired lists follow network order),
 * nemo_apply_stdp applies to every network, nemo_delete_simulation frees it.
 * Single-network queries address network 0, and nemo_step is refused for
 * more than one network.  No counterpart in the reference. */
nemo_status_t nemo_hip_new_batch(const nemo_network_t nets[], size_t count, nemo_configuration_t,
		nemo_simulation_t* sim);
nemo_status_t nemo_hip_batch_size(nemo_simulation_t, unsigned* count);
nemo_status_t nemo_hip_batch_get_neuron_state(nemo_simulation_t, unsigned net, unsigned neuron, unsigned var,
		float* val);
nemo_status_t nemo_hip_batch_get_neuron_parameter(nemo_simulation_t, unsigned net, unsigned neuron,
		unsigned param, float* val);
nemo_status_t nemo_hip_batch_get_membrane_potential(nemo_simulation_t, unsigned net, unsigned neuron, float* val);

/* Extension: capture the step graphs nemo_hip_run(nsteps, stdp_period) uses,
 * so that a timed run excludes graph construction. */
nemo_status_t nemo_hip_prepare(nemo_simulation_t, unsigned nsteps, unsigned stdp_period);

/* Extension: launch shape of the per-step kernels, NEMO_HIP_SHAPE_FIELDS
 * values: shape[0] delivery blocks for the delays 2..D (for all delays when
 * shape[1] is 0), shape[1] delivery blocks for the delay 1 (both together at
 * most 1024), shape[2] events a delivery block stages in LDS (64..32768),
 * shape[3] where the STDP accumulation runs: 0 a launch of its own on a graph
 * branch beside the delivery, 1 inside the neuron-update kernels, 2 in the
 * delivery launch that follows the previous update, 3 a launch of its own on
 * a graph branch that may run until the next step's update, 4 in one launch
 * for up to 64 steps after them (every run of steps ends with it); shape[4] firing-bitmap
 * words per STDP block (1..8, 1..64 in STDP mode 4), shape[5] when split, how the delays 2..D of the next step are
 * delivered beside the current update within nemo_hip_run: 0 not at all, 1 on
 * a graph branch, 2 as extra blocks of the update launch (which then also
 * carries the STDP blocks of STDP mode 2), 3 as in 2 with the update's tiles
 * delivering the delay-1 events of their own firing too, so that every step
 * of a run after the first is one launch.  Every shape gives identical
 * results. */
#define NEMO_HIP_SHAPE_FIELDS 6
nemo_status_t nemo_hip_get_launch_shape(nemo_simulation_t, unsigned shape[NEMO_HIP_SHAPE_FIELDS]);
nemo_status_t nemo_hip_set_launch_shape(nemo_simulation_t, const unsigned shape[NEMO_HIP_SHAPE_FIELDS]);

/* Extension: autotuner.  Times candidate launch shapes (coordinate descent
 * over the shape parameters), each over steps_per_trial steps from the same
 * snapshot of the simulation state, then restores that state and keeps the
 * fastest shape, returned in shape (may be NULL).  Results of the simulation
 * are unaffected. */
nemo_status_t nemo_hip_autotune(nemo_simulation_t, unsigned steps_per_trial, unsigned shape[NEMO_HIP_SHAPE_FIELDS]);

/* Extension: float-weight mode with a hand-written adjoint (no counterpart in
 * the reference, whose only mode is the fixed-point one).  A network of one
 * Izhikevich type with neuron indices 0..N-1, simulated with FP32 weights and
 * FP32 current sums; where spikes feed synapses and the loss, the threshold
 * is differentiated through the fast-sigmoid surrogate
 * dS/dv = 1 / (1 + 10 |v - 30|)^2 (the reset and the substep stops are held
 * constant).  Per step t:
 *   I = x[t] + sum of w over synapses whose source spiked at t - delay;
 *   four 0.25 ms substeps of the Izhikevich step, each only while v < 30
 *   (v <- v + h (0.04 v^2 + 5 v + 140 - u + I), u <- u + h a (b v - u));
 *   S[t] = [v >= 30]; v <- v + S (c - v); u <- u + S d; V[t] = v.
 * All arrays are device pointers owned by the caller: weights w in network
 * order (source ascending, then insertion order, as synapse ids), per-neuron
 * a, b, c, d, v0, u0 by index, x, spikes and v_out [steps][N] row-major;
 * checkpoints, double [ceil(steps / checkpoint_every)][2][N], holds (v, u)
 * before every checkpoint_every-th step (the state is kept in double; the
 * weights and their per-step sum are float).  Work is queued on `stream` (a
 * hipStream_t; NULL for the default stream); calls on one handle must not
 * overlap. */
typedef struct nemo_hip_float_network_class* nemo_hip_float_network_t;
nemo_status_t nemo_hip_new_float_network(nemo_network_t, int device, nemo_hip_float_network_t* out);
void nemo_hip_delete_float_network(nemo_hip_float_network_t);
nemo_status_t nemo_hip_float_network_sizes(nemo_hip_float_network_t, size_t* neurons, size_t* synapses,
		unsigned* max_delay);
nemo_status_t nemo_hip_float_forward(nemo_hip_float_network_t, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* v0, const float* u0, const float* x,
		float* spikes, float* v_out, double* checkpoints);
/* Extension (no counterpart in the reference): nemo_hip_float_forward for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results; the forward's workspace -- weights in reverse order (4 B per
 * synapse), the state in double (16 B per neuron) and the spike bitmaps
 * (4 B per 32 neurons, rounded up to 16 B, per step) -- is carved from a
 * static device array of 64 MiB per device instead of the handle's own
 * allocations, and nothing is allocated or freed.  Fails with
 * NEMO_INVALID_INPUT when the workspace for `steps` steps exceeds it.  Every
 * call on one device shares that array (it keeps the last call's contents,
 * none of which a call reads before writing): calls that overlap on one
 * device race. */
nemo_status_t nemo_hip_float_forward_static(nemo_hip_float_network_t, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* v0, const float* u0, const float* x,
		float* spikes, float* v_out, double* checkpoints);
/* The adjoint of nemo_hip_float_forward: from dL/dspikes and dL/dv_out (either
 * may be NULL) and the forward's spikes and checkpoints, the gradients of the
 * loss with respect to w, a, b, d and x (grad_x may be NULL) and the initial
 * state v0, u0 (either may be NULL).  Each checkpoint segment is replayed
 * forward from its checkpoint and the recorded spikes (reproducing the
 * forward's state exactly), then stepped backward in time: the transposed
 * delivery gathers each source's spike adjoint from the current adjoints of
 * its targets at step + delay. */
nemo_status_t nemo_hip_float_backward(nemo_hip_float_network_t, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* x, const float* spikes, const double* checkpoints,
		const float* grad_spikes, const float* grad_v, float* grad_w, float* grad_a, float* grad_b,
		float* grad_d, float* grad_x, float* grad_v0, float* grad_u0);

/* Extension (parity inspection): float currents (E then I) delivered to each
 * neuron in the most recent step, in simulation order.  The reference CPU
 * backend holds these in m_currentE / m_currentI (cpu/Simulation.cpp:279-288).
 * Requires nemo_hip_set_debug(sim, 1) before the step. */
nemo_status_t nemo_hip_set_debug(nemo_simulation_t, int enable);
nemo_status_t nemo_hip_get_currents(nemo_simulation_t, float E[], float I[], size_t n);
/* Extension (parity inspection): simulation-order -> user index map */
nemo_status_t nemo_hip_local_to_global(nemo_simulation_t, unsigned gidx[], size_t n);
/* Extension (parity inspection): 64-bit firing history, simulation order
 * (cpu/Simulation.hpp m_recentFiring) */
nemo_status_t nemo_hip_get_recent_firing(nemo_simulation_t, uint64_t hist[], size_t n);
/* Extension (parity inspection): RNG state, 4 words per neuron, simulation order */
nemo_status_t nemo_hip_get_rng_state(nemo_simulation_t, uint32_t state[], size_t n);
/* Extension (parity inspection): state variable var of every neuron in
 * simulation order (NaN for neurons whose model has fewer state variables);
 * one copy instead of one nemo_get_neuron_state_s call per neuron
 * (src/nemo.h:437-446). */
nemo_status_t nemo_hip_get_neuron_states(nemo_simulation_t, unsigned var, float out[], size_t n);
/* Extension (parity inspection): Q11.20 weight and STDP accumulator of every
 * synapse in network order (source ascending, then insertion order) */
nemo_status_t nemo_hip_get_fixed_weights(nemo_simulation_t, int32_t w[], size_t n);
nemo_status_t nemo_hip_get_stdp_accumulators(nemo_simulation_t, int32_t acc[], size_t n);
/* Extension: synapse and neuron counts of a built simulation */
nemo_status_t nemo_hip_counts(nemo_simulation_t, size_t* neurons, size_t* synapses, size_t* plastic);

/* Extension (measurement): run nsteps steps (no stimulus) with HIP events
 * around each phase on the simulation's stream.  avg_us receives the mean
 * duration of the propagate launch (delivery and STDP accumulation) over the
 * steps that have one, of the neuron updates (all models; with launch-shape
 * overlap 2 or 3 the update launch also delivers the next step) per step,
 * the propagate launches per step (below 1 with overlap 3), and a zero
 * (reserved); counters the events delivered, neurons fired, plastic synapses
 * examined and steps run.  The
 * reference's Timer (src/nemo/Timer.hpp:23-79) only measures whole steps. */
nemo_status_t nemo_hip_profile(nemo_simulation_t, unsigned nsteps, double avg_us[4], uint64_t counters[4]);

/* Extension (diagnostics): evaluate the device's glibc-exact logf (fn 0),
 * sinf (fn 1) or expf (fn 2) on n host inputs, results to host out.  The
 * reference calls the host C library for these (RNG.cpp:28-32, HH.cpp:92-98). */
nemo_status_t nemo_hip_eval_math(int fn, const float in[], float out[], size_t n);

/* Brian's propagate (src/nemo/Simulation.hpp:93-113; cpu/Simulation.cpp:
 * 115-137 and cuda/Simulation.cpp:328-346, which NeMo compiles with
 * NEMO_BRIAN_ENABLED): instead of updating the neurons, the firing of the
 * next step is given -- the nfired neuron indices in fired -- its spikes are
 * delivered, and E and I (n = neuron count entries each, simulation order as
 * nemo_hip_local_to_global lists it) receive the excitatory and inhibitory
 * synaptic input of the step after it, as the neurons would see it.  Neuron
 * state is untouched; the simulation advances one step.  Refused when STDP
 * is configured (as in the CUDA backend) and for batches. */
nemo_status_t nemo_hip_propagate(nemo_simulation_t, const unsigned fired[], size_t nfired, float E[], float I[],
		size_t n);
/* The device form (the CUDA backend's): fired is a device array of one
 * uint32 flag per neuron in simulation order (nonzero: fired); *E and *I
 * receive device arrays of the input in simulation order, valid until the
 * next call on the simulation.  The work is complete on return. */
nemo_status_t nemo_hip_propagate_device(nemo_simulation_t, const uint32_t* fired, float** E, float** I);

/* --------------------------------------------------------- backend factory */

/* NeMo resolves a backend's factory by name (src/nemo/nemo.cpp:92-101,
 * 120-133): cuda_simulation, cuda_set_configuration, cuda_test_device,
 * cuda_device_count and cuda_device_description for the CUDA backend
 * (src/nemo/cuda/create_simulation.hpp:14-22, src/nemo/cuda/devices.hpp:
 * 21-56), whose arguments are NeMo's own C++ objects.  Their HIP
 * counterparts, hip_simulation and the rest with the same signatures, are
 * defined by integration/nemo_hip_backend.cpp, which NeMo compiles against
 * its headers and which implements nemo::SimulationBackend over this C ABI.
 * This library provides the device test they need: */

/* cuda_test_device (src/nemo/cuda/devices.hpp:33-37), as a status: the
 * device exists and accepts an allocation */
nemo_status_t nemo_hip_test_device(unsigned device);

#ifdef __cplusplus
}
#endif

#endif
