/*
 * Launch interface of the backend's own kernels (kernels.hip).  Plain C ABI:
 * device pointers, sizes and a stream; each returns the launch status.  These
 * replace the CUDA backend's launch wrappers (src/nemo/cuda/kernel.hpp:35-116).
 */
#ifndef NEMO_AMD_KERNELS_H
#define NEMO_AMD_KERNELS_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "nemo_hip_plugin.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Longest supported synaptic delay (ms); the firing-history word is 64 bits. */
#define NEMO_MAX_DELAY 64
/* Steps remembered in the step-start ring (power of two, > max delay + 1). */
#define NEMO_STEP_RING 128

/* Spike delivery: for every delay d in 1..D and every neuron n that fired at
 * step t-d, add the weights of the synapse row (n, d) into the target
 * accumulators, excitatory (w >= 0) and inhibitory separately.
 * Semantics of cpu::Simulation::deliverSpikes (cpu/Simulation.cpp:259-308). */
typedef struct nemo_deliver_args {
	nemo_hip_clock* clock;           /* reads step, adds delivered events */
	const uint64_t* step_start;      /* ring of NEMO_STEP_RING log positions */
	const uint32_t* fired_log;
	uint64_t log_mask;
	const uint32_t* rowptr;          /* N*D+1 offsets into fcm, row (n,d) = n*D + d-1 */
	const int2* fcm;                 /* {target simulation index, Q11.20 weight} */
	int64_t* acc;                    /* E plane, then I plane at acc_pitch */
	size_t acc_pitch;
	unsigned D;                      /* rows per source (network max delay) */
} nemo_deliver_args;

hipError_t nemo_launch_deliver(hipStream_t, const nemo_deliver_args*, unsigned lanes_per_row, unsigned blocks);

/* STDP accumulation for the neurons that fired P = |postfire| steps ago
 * (ConnectivityMatrix::accumulateStdp, ConnectivityMatrix.cpp:224-258,
 * StdpProcess::weightChange, StdpProcess.cpp:68-124). */
typedef struct nemo_stdp_args {
	nemo_hip_clock* clock;           /* reads step, adds examined synapses */
	const uint64_t* step_start;
	const uint32_t* fired_log;
	uint64_t log_mask;
	const uint32_t* rcm_ptr;         /* N+1 offsets of each target's plastic incoming synapses */
	const uint32_t* rcm_sd;          /* source simulation index | (delay-1) << 26 */
	int32_t* rcm_acc;                /* accumulated weight change */
	const uint64_t* hist;
	unsigned postPreWindow;          /* P */
	uint64_t potentiationBits, depressionBits, prePostBits, postPreBits;
	int32_t fnPre[64];
	int32_t fnPost[64];
} nemo_stdp_args;

hipError_t nemo_launch_stdp(hipStream_t, const nemo_stdp_args*, unsigned lanes_per_row, unsigned blocks);

/* applyStdp (ConnectivityMatrix::applyStdp, ConnectivityMatrix.cpp:272-325 and
 * StdpProcess::updatedWeight, StdpProcess.cpp:128-138). */
typedef struct nemo_apply_args {
	const uint32_t* rcm_fwd;         /* fcm index of each plastic synapse */
	int32_t* rcm_acc;
	int2* fcm;
	size_t count;
	int32_t reward;                  /* Q11.20 reward; 0 clears without adding */
	int32_t minE, maxE, minI, maxI;
	unsigned fbits;
} nemo_apply_args;

hipError_t nemo_launch_apply_stdp(hipStream_t, const nemo_apply_args*);

/* End of step: step += 1, cycle += 1, step_start[step] = log_head. */
hipError_t nemo_launch_advance(hipStream_t, nemo_hip_clock*, uint64_t* step_start);

/* Firing stimulus: set (or clear) the bits of the n given simulation indices. */
hipError_t nemo_launch_fstim(hipStream_t, uint32_t* bits, const uint32_t* idx, size_t n, int set);

/* Current stimulus: ext[idx[i]] = val[i] (indices must be distinct). */
hipError_t nemo_launch_istim(hipStream_t, float* ext, const uint32_t* idx, const float* val, size_t n);

/* Device logf (fn 0), sinf (fn 1) or expf (fn 2) of n inputs (diagnostics). */
hipError_t nemo_launch_eval_math(hipStream_t, int fn, const float* in, float* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
