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
 * update kernels gather the bins), in two launches over two block sets: the
 * delays 2..D (which only need firing up to step t-2, so this launch may run
 * beside the update of step t-1) and the delay 1.  For every delay d of the
 * set and every neuron n that fired at step t-d, the synapse row (n, d) is
 * delivered: each event {target, w} is counting-sorted in LDS by target tile
 * and written to the block's bin region, bin_table[t % 2][tile][column] =
 * count << 16 | offset, column = block_base + block.  Events beyond the
 * block's bin capacity are added with global atomics into acc plane t % 2 (E
 * if w >= 0, else I).  Block b of a set covers the (source tile, delay) pairs
 * q in [pair_bounds[b], pair_bounds[b+1]), q = tile * d_count + d - d_first,
 * and every target of those rows lies in its tile window [windows[b].x,
 * windows[b].y); it writes the table entries of exactly the tiles of its
 * window.  Semantics of cpu::Simulation::deliverSpikes (cpu/Simulation.cpp:
 * 259-308).
 *
 * STDP accumulation of step t-1 for the neurons that fired at step t-1-P
 * (the contract of nemo_hip_stdp) may share the launch that delivers step t
 * after the update of step t-1 (blocks after the delivery blocks), run as a
 * launch of its own (stdp_kernel), or be left to the update kernels.  It
 * reads the firing history after step t-1, plane (t-1) % 2 of hist.  Skipped
 * when t == 0 or when step t-1's accumulation already ran
 * (clock->stdp_flushed == t-1). */
typedef struct nemo_propagate_args {
	const nemo_hip_clock* clock;
	uint64_t step_offset;            /* t = clock->base_step + step_offset */
	const uint64_t* fired_bits;      /* NEMO_STEP_RING bitmaps of fired_words words */
	size_t fired_words;
	/* delivery: one block set */
	unsigned deliver_blocks;         /* blocks of the set */
	unsigned block_base;             /* table column (and bin region) of its first block */
	unsigned table_columns;          /* blocks of both sets */
	unsigned d_first, d_count;       /* delays of the set */
	const uint32_t* pair_bounds;     /* [deliver_blocks+1] first pair of each block */
	const uint2* windows;            /* [deliver_blocks] target tiles [x, y) of each block */
	unsigned max_window;             /* largest window (tiles) */
	const uint32_t* rowptr;          /* (neurons*D)+1 offsets, row (n,d) = n*D + d-1 */
	const int2* fcm;                 /* {target device index, Q11.20 weight} */
	unsigned D;                      /* rows per source (network max delay) */
	unsigned tiles;
	unsigned bin_capacity;           /* events per block region */
	uint32_t* bin_table;             /* [2][tiles][table_columns] */
	int2* bins;                      /* [2][table_columns][bin_capacity] {target % NEMO_TILE, w} */
	int64_t* acc;                    /* [2][2][acc_pitch] overflow accumulators, E plane then I plane */
	size_t acc_pitch;
	unsigned long long* block_events; /* [table_columns] events delivered, accumulated */
	/* STDP in its own launch */
	unsigned stdp_words;             /* bitmap words per STDP block (1..4) */
	const uint64_t* hist;            /* two planes of hist_pitch */
	size_t hist_pitch;
	nemo_hip_stdp stdp;              /* examined: one counter per STDP block */
} nemo_propagate_args;

/* dynamic LDS bytes a delivery block needs */
size_t nemo_propagate_lds_bytes(unsigned max_window, unsigned bin_capacity);
/* Delivery (a->deliver_blocks blocks; none if 0) and STDP accumulation
 * (stdp_blocks blocks; none if 0): one launch with the STDP blocks after the
 * delivery blocks when fuse_stdp is set, else the STDP launch first. */
hipError_t nemo_launch_propagate(hipStream_t, const nemo_propagate_args*, unsigned stdp_blocks, int fuse_stdp);

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
