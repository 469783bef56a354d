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
/* Threads per block of the backend's kernels. */
#define NEMO_BLOCK 256

/* Per step, before the neuron updates:
 *
 * Spike delivery of step t (propagate_kernel), binned by target tile (the
 * update kernels gather the bins).  For every delay d in 1..D and every
 * neuron n that fired at step t-d, the synapse row (n, d) is delivered: each
 * event {target, w} is counting-sorted in LDS by target tile and written to
 * the block's bin region; bin_table[tile][block] = count << 16 | offset.
 * Events beyond the block's bin capacity are added with global atomics into
 * acc (E if w >= 0, else I).  Delivery block b covers the (source tile,
 * delay) pairs q in [pair_bounds[b], pair_bounds[b+1]), q = tile*D + d-1, and
 * every target of those rows lies in its tile window [windows[b].x,
 * windows[b].y); it writes the table entries of exactly the tiles of its
 * window.  Semantics of cpu::Simulation::deliverSpikes (cpu/Simulation.cpp:
 * 259-308).
 *
 * Optionally, before delivery, STDP accumulation of step t-1 in a launch of
 * its own (stdp_kernel), for the neurons that fired at step t-1-P (the
 * contract of nemo_hip_stdp); usually the update kernels do it instead.  It
 * reads the firing history after step t-1, plane (t-1) % 2 of hist.  Skipped
 * when t == 0 or when step t-1's accumulation already ran
 * (clock->stdp_flushed == t-1). */
typedef struct nemo_propagate_args {
	const nemo_hip_clock* clock;
	uint64_t step_offset;            /* t = clock->base_step + step_offset */
	const uint64_t* fired_bits;      /* NEMO_STEP_RING bitmaps of fired_words words */
	size_t fired_words;
	/* delivery */
	unsigned deliver_blocks;
	const uint32_t* pair_bounds;     /* [deliver_blocks+1] first pair of each block */
	const uint2* windows;            /* [deliver_blocks] target tiles [x, y) of each block */
	unsigned max_window;             /* largest window (tiles) */
	const uint32_t* rowptr;          /* (neurons*D)+1 offsets, row (n,d) = n*D + d-1 */
	const int2* fcm;                 /* {target device index, Q11.20 weight} */
	unsigned D;                      /* rows per source (network max delay) */
	unsigned tiles;
	unsigned bin_capacity;           /* events per block region */
	uint32_t* bin_table;             /* [tiles][deliver_blocks] */
	int2* bins;                      /* [deliver_blocks][bin_capacity] {target % NEMO_TILE, w} */
	int64_t* acc;                    /* overflow accumulators, E plane then I plane */
	size_t acc_pitch;
	unsigned long long* block_events; /* [deliver_blocks] events delivered, accumulated */
	/* STDP in its own launch */
	unsigned stdp_words;             /* bitmap words per STDP block (1..4) */
	const uint64_t* hist;            /* two planes of hist_pitch */
	size_t hist_pitch;
	nemo_hip_stdp stdp;              /* examined: one counter per STDP block */
} nemo_propagate_args;

/* dynamic LDS bytes a delivery block needs */
size_t nemo_propagate_lds_bytes(unsigned max_window, unsigned bin_capacity);
/* STDP accumulation (stdp_blocks blocks; none if 0), then delivery
 * (a->deliver_blocks blocks; none if 0). */
hipError_t nemo_launch_propagate(hipStream_t, const nemo_propagate_args*, unsigned stdp_blocks);

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

/* Firing stimulus: set (or clear) the bits of the n given device indices. */
hipError_t nemo_launch_fstim(hipStream_t, uint32_t* bits, const uint32_t* idx, size_t n, int set);

/* Current stimulus: ext[idx[i]] = val[i] (indices must be distinct). */
hipError_t nemo_launch_istim(hipStream_t, float* ext, const uint32_t* idx, const float* val, size_t n);

/* Plain device-to-device copy of bytes (a multiple of 16) (measurement). */
hipError_t nemo_launch_copy(hipStream_t, const void* src, void* dst, size_t bytes);

/* Device logf (fn 0), sinf (fn 1) or expf (fn 2) of n inputs (diagnostics). */
hipError_t nemo_launch_eval_math(hipStream_t, int fn, const float* in, float* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
