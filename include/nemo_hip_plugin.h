/*
 * nemo_hip_plugin.h -- ABI between the HIP backend and its neuron-model
 * plugins.
 *
 * Each neuron model is a shared library <plugin dir>/hip/lib<Model>.so that
 * exports hip_update_neurons and hip_init_neurons, loaded by model name the
 * way NeMo loads <plugin dir>/cuda/lib<Model>.so (src/nemo/cuda/Neurons.cpp:
 * 43-44, src/nemo/Plugin.cpp:141-180).  The two entry points replace
 * cuda_update_neurons / cuda_init_neurons (src/nemo/cuda/plugins/
 * neuron_model.h:26-62); the model's .ini file gives its parameter and state
 * counts as in NeMo (src/nemo/NeuronType.cpp:93-158).
 *
 * Contract of hip_update_neurons, for neurons [start, start+count) of one type
 * in the current step (the semantics of cpu_update_neurons,
 * src/nemo/cpu/plugins/neuron_model.h:25-51):
 *   - read and clear the accumulated Q(31-fbits).fbits currents acc[E plane]
 *     and acc[I plane], saturating each to the int32 range before converting
 *     to float (src/nemo/fixedpoint.cpp:40-49);
 *   - read and clear ext_current when it is not NULL;
 *   - apply one 1 ms step of the model; a neuron fires if the model says so or
 *     if its bit in fstim is set (fstim may be NULL);
 *   - shift the neuron's 64-bit firing history left by one, OR-ing in firing;
 *   - append each fired neuron's simulation index to the fired log at
 *     position (clock->log_head++ & log_mask), in any order.
 * All pointers are device pointers; the launch goes on the given stream and
 * must not synchronise it, so that the backend can capture it in a graph.
 */
#ifndef NEMO_HIP_PLUGIN_H
#define NEMO_HIP_PLUGIN_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device-resident step clock, advanced by the backend at the end of a step. */
typedef struct nemo_hip_clock {
	uint64_t step;     /* steps completed since construction (monotonic) */
	uint64_t cycle;    /* simulated ms since the last timer reset (Timer, src/nemo/Timer.hpp:23-79) */
	uint64_t log_head; /* fired-log entries appended since construction */
	uint64_t events;   /* synaptic events delivered since construction */
	uint64_t stdp_synapses; /* plastic synapses examined by STDP accumulation */
} nemo_hip_clock;

/* One neuron type's state: structure of arrays, plane p of a quantity starts
 * at element p*pitch. */
typedef struct nemo_hip_neuron_group {
	unsigned start;    /* simulation index of the first neuron */
	unsigned count;    /* neurons in the group */
	size_t pitch;      /* elements between planes */
	float* param;      /* nParam planes */
	float* state;      /* nState planes (history length 1) */
	uint32_t* rng;     /* 4 xorshift words per neuron, interleaved; may be NULL */
} nemo_hip_neuron_group;

/* Per-simulation buffers shared by all groups in a step. */
typedef struct nemo_hip_step_buffers {
	int64_t* acc;          /* accumulated currents: E plane [0, acc_pitch), I plane after it */
	size_t acc_pitch;
	float* ext_current;    /* per simulation index, or NULL */
	const uint32_t* fstim; /* firing-stimulus bit-vector, or NULL */
	uint64_t* hist;        /* 64-bit firing history per simulation index */
	uint32_t* fired_log;   /* circular log of fired simulation indices */
	uint64_t log_mask;     /* log capacity - 1 (capacity is a power of two) */
	nemo_hip_clock* clock;
	unsigned fbits;        /* fractional bits of the fixed-point currents */
	float* debug_current;  /* NULL, or 2 planes of acc_pitch floats receiving E and I */
} nemo_hip_step_buffers;

typedef hipError_t hip_update_neurons_t(hipStream_t stream,
		const nemo_hip_neuron_group* group, const nemo_hip_step_buffers* buffers);

typedef hipError_t hip_init_neurons_t(hipStream_t stream,
		const nemo_hip_neuron_group* group, const nemo_hip_step_buffers* buffers);

#ifdef __cplusplus
}
#endif

#endif
