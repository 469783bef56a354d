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
/* Threads per block of the backend's small kernels. */
#define NEMO_BLOCK 256
/* Threads per block of the propagate launches (delivery and STDP blocks):
 * one block per CU at a time does most of a step's delivery, so it gets
 * several waves per SIMD to hide LDS and memory latency. */
#ifndef NEMO_PROPAGATE_BLOCK
#define NEMO_PROPAGATE_BLOCK 512
#endif
/* Firing-bitmap words (64 targets each) an STDP block may cover: one target
 * per thread, or, in STDP batches, groups of that many in turn. */
#define NEMO_STDP_MAX_WORDS (NEMO_PROPAGATE_BLOCK / 64)
#define NEMO_STDP_BATCH_MAX_WORDS 64
/* Threads per STDP batch block, and the waves per SIMD its registers allow
 * (80 VGPRs): one wave per SIMD, which fits beside three blocks of an update
 * launch that carries delivery (NEMO_COLAUNCH_WAVES), so that a batch running
 * beside the steps does not take their blocks' places. */
#define NEMO_STDP_BATCH_BLOCK 256
#define NEMO_STDP_BATCH_WAVES 6
/* Default bitmap words per batch block: 261 blocks for 100k neurons, about
 * one per CU (100k x 1000 with STDP, us per step in one box: 4 words 23.2,
 * 5 words 22.4, 6 words 22.1, 7 words 22.2, 8 words 22.6) */
#define NEMO_STDP_BATCH_WORDS 6


/* dynamic LDS bytes a delivery block needs */
size_t nemo_propagate_lds_bytes(unsigned max_window, unsigned bin_capacity);
/* The pending STDP accumulation (clock->stdp_first, stdp_count) in one launch
 * of stdp_blocks blocks, from the two history snapshots.  It reads neither
 * the live history nor the bitmaps, so it may run beside later steps. */
hipError_t nemo_launch_stdp_batch(hipStream_t, const nemo_propagate_args*, unsigned stdp_blocks);
/* The same as a flush: afterwards nothing is pending and clock->stdp_flushed
 * is step base + offset (set by a kernel after the batch, so that no copy
 * from the host follows it) */
hipError_t nemo_launch_stdp_flush(hipStream_t, const nemo_propagate_args*, unsigned stdp_blocks, uint64_t offset);
/* A snapshot of the history after step base + offset (at most 64 steps after
 * the last one); the steps since the last snapshot become the pending
 * accumulation.  The pending batch must have run.  Then the clock's step
 * base moves on by advance (0: it stays). */
hipError_t nemo_launch_stdp_snapshot(hipStream_t, const nemo_propagate_args*, uint64_t offset, uint64_t advance);
/* dynamic LDS an update launch carrying propagate blocks needs (delivery
 * blocks when delivery is set, else STDP blocks only; the launch's update
 * blocks share the buffer) */
size_t nemo_colaunch_lds_bytes(unsigned max_window, unsigned bin_capacity, int delivery);
/* Static LDS of the propagate kernel (bytes). */
size_t nemo_propagate_static_lds(void);
/* Delivery (a->deliver_blocks blocks; none if 0) and STDP accumulation
 * (stdp_blocks blocks; none if 0): one launch with the STDP blocks after the
 * delivery blocks when fuse_stdp is set, else the STDP launch first.  start
 * and stop (may be NULL) are recorded at the start of the first kernel and
 * the end of the last. */
hipError_t nemo_launch_propagate(hipStream_t, const nemo_propagate_args*, unsigned stdp_blocks, int fuse_stdp,
		hipEvent_t start, hipEvent_t stop);

/* Both delivery sets of one step (a's blocks, then a2's; the same step,
 * clock, bin capacity, window and table) in one launch, with stdp_blocks STDP
 * blocks of a's arguments after them. */
hipError_t nemo_launch_propagate_sets(hipStream_t, const nemo_propagate_args* a, const nemo_propagate_args* a2,
		unsigned stdp_blocks, hipEvent_t start, hipEvent_t stop);

/* applyStdp (ConnectivityMatrix::applyStdp, ConnectivityMatrix.cpp:272-325 and
 * StdpProcess::updatedWeight, StdpProcess.cpp:128-138). */
typedef struct nemo_apply_args {
	const uint32_t* rcm_of;          /* per FCM synapse: its accumulator index, or ~0 if static */
	int32_t* rcm_acc;
	int2* fcm;
	size_t count;                    /* plastic synapses */
	size_t fcm_count;                /* FCM synapses */
	int32_t reward;                  /* Q11.20 reward; 0 clears without adding */
	int32_t minE, maxE, minI, maxI;
	unsigned fbits;
} nemo_apply_args;

hipError_t nemo_launch_apply_stdp(hipStream_t, const nemo_apply_args*);
/* rcm_of (fcm_count entries): the accumulator index of every plastic
 * synapse, ~0 for the others, from rcm_fwd (count entries). */
hipError_t nemo_launch_invert_rcm(hipStream_t, const uint32_t* rcm_fwd, uint32_t* rcm_of, size_t count,
		size_t fcm_count);

/* Firing of step base_step given from outside (one flag per neuron in
 * simulation order; dev_to_sim maps device index -> simulation index, ~0
 * for padding): histories and the step's bitmap written as by an update.
 * One block per tile. */
hipError_t nemo_launch_external_fire(hipStream_t, const nemo_hip_step_buffers* b, const uint32_t* flags,
		const uint32_t* dev_to_sim, unsigned tiles);
/* Synaptic input of step base_step gathered per tile (overflow accumulators
 * cleared) and written as floats in simulation order. */
hipError_t nemo_launch_gather_currents(hipStream_t, const nemo_hip_step_buffers* b, const uint32_t* dev_to_sim,
		float* E, float* I, unsigned tiles);

/* Firing readout of k steps from step t0 (their firing-bitmap ring slots,
 * words words each): counts[s] the neurons fired in step t0 + s, and list
 * their l2g entries, step after step, each step in ascending device index
 * (list holds up to k * words * 64 entries). */
hipError_t nemo_launch_fired_list(hipStream_t, const uint64_t* bits, size_t words, uint64_t t0, unsigned k,
		const uint32_t* l2g, uint32_t* counts, uint32_t* list);

/* Firing stimulus: set (or clear) the bits of the n given device indices. */
hipError_t nemo_launch_fstim(hipStream_t, uint32_t* bits, const uint32_t* idx, size_t n, int set);

/* Current stimulus: ext[idx[i]] = val[i] (indices must be distinct). */
hipError_t nemo_launch_istim(hipStream_t, float* ext, const uint32_t* idx, const float* val, size_t n);

/* Device logf (fn 0), sinf (fn 1) or expf (fn 2) of n inputs (diagnostics). */
hipError_t nemo_launch_eval_math(hipStream_t, int fn, const float* in, float* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
