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
 * Neurons live in the backend's device index space: each neuron type occupies
 * a range starting at a multiple of NEMO_TILE, and neuron i of the type has
 * device index start + i.  One update block of NEMO_TILE threads owns one
 * tile of NEMO_TILE consecutive device indices.
 *
 * Contract of hip_update_neurons, for neurons [start, start+count) of one type
 * in step s = clock->base_step + step_offset (the semantics of
 * cpu_update_neurons, src/nemo/cpu/plugins/neuron_model.h:25-51):
 *   - sum the step's synaptic input of the tile: the binned events of the
 *     delivery blocks in the tile's two bin ranges (bin_table / bins, the
 *     buffers of step parity s % 2; the first set's only when bin_first_set
 *     is nonzero) plus the overflow accumulators acc (of the same parity),
 *     excitatory (w >= 0) and inhibitory separately, saturate each sum to the
 *     int32 range and convert to float (src/nemo/fixedpoint.cpp:40-49); clear
 *     the overflow accumulators;
 *   - read and clear ext_current when it is not NULL;
 *   - apply one 1 ms step of the model; a neuron fires if the model says so or
 *     if its bit in fstim is set (fstim may be NULL);
 *   - shift the neuron's 64-bit firing history left by one, OR-ing in firing:
 *     the history of neuron n after step s is
 *     hist[NEMO_HIST_INDEX(n, s, hist_pitch)] (plane s % 2), so the update
 *     reads the word of step s-1 and writes the word of step s;
 *   - write the tile's words of the step's firing bitmap,
 *     fired_bits[(s % NEMO_STEP_RING) * fired_words + index / 64];
 *   - when co.push is set, deliver the delay-1 events of the step's firing
 *     for step s+1 (see nemo_hip_colaunch);
 *   - when stdp.rcm_ptr is not NULL, run the STDP accumulation of step s-1 for
 *     the tile's neurons (see nemo_hip_stdp).
 * The helpers in nemo_amd/csrc/plugin_device.h implement all of this but the
 * model step.  All pointers are device pointers; the launch goes on the given
 * stream and must not synchronise it, so that the backend can capture it in a
 * graph.
 */
#ifndef NEMO_HIP_PLUGIN_H
#define NEMO_HIP_PLUGIN_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Neurons per tile (and threads per update block). */
#define NEMO_TILE 256
/* Steps of firing bitmaps kept (power of two, > maximum delay + STDP window). */
#define NEMO_STEP_RING 128
/* Threads per block of an update launch that carries propagate blocks
 * (nemo_hip_colaunch): the propagate blocks use all of them, a tile's block
 * its first NEMO_TILE (one per neuron; the others end at once, or first help
 * to gather the tile's input). */
#define NEMO_COLAUNCH_BLOCK 512
/* Waves per SIMD an update kernel compiled for such launches allows
 * (__launch_bounds__(NEMO_COLAUNCH_BLOCK, NEMO_COLAUNCH_WAVES): at most 72
 * VGPRs), so that a CU holds three of its blocks and, beside them, one wave
 * per SIMD of an STDP batch running beside the steps. */
#define NEMO_COLAUNCH_WAVES 7
/* Most delivery blocks (bin regions) a step's input can come from. */
#define NEMO_MAX_BIN_BLOCKS 1024
/* Firing history: two planes of hist_pitch words, plane s % 2 after step s. */
#define NEMO_HIST_INDEX(n, s, pitch) (((s) & 1) * size_t(pitch) + size_t(n))

/* Device-resident step base, set by the host before launches. */
typedef struct nemo_hip_clock {
	uint64_t base_step;    /* added to each launch's step_offset */
	uint64_t cycle_origin; /* step of the last timer reset: cycle = step - cycle_origin
	                          (Timer, src/nemo/Timer.hpp:23-79) */
	uint64_t stdp_flushed; /* step whose STDP accumulation already ran on its own (~0: none) */
	/* STDP batches: snapshots of the firing history taken so far (the last
	 * in slot stdp_snaps % 2, the one before in the other slot), the step of
	 * the last, and the steps [stdp_first, stdp_first + stdp_count) between
	 * the two whose accumulation is pending (stdp_count 0: none) */
	uint64_t stdp_snaps;
	uint64_t stdp_last;
	uint64_t stdp_first;
	uint64_t stdp_count;
} nemo_hip_clock;

/* One neuron type's state: structure of arrays, plane p of a quantity starts
 * at element p*pitch. */
typedef struct nemo_hip_neuron_group {
	unsigned start;    /* device index of the first neuron (multiple of NEMO_TILE) */
	unsigned count;    /* neurons in the group */
	size_t pitch;      /* elements between planes */
	float* param;      /* nParam planes */
	float* state;      /* nState planes (history length 1) */
	uint32_t* rng;     /* 4 xorshift words per neuron, interleaved; may be NULL */
	float* scratch;    /* the planes the model asks for with scratch-planes under [hip] in its .ini
	                      (pitch apart), owned by the simulation; NULL if none.  Contents are the
	                      model's own between its launches of a step; nothing else reads them */
	unsigned history;  /* state history length H (the .ini's history): state holds H slots of
	                      nState planes, slot h plane v at (h * nState + v) * pitch.  With H > 1
	                      the update of step s reads slot c % H and writes slot (c + 1) % H, c =
	                      s - clock->cycle_origin (cpu/Neurons.cpp:123; the reference's 32-bit
	                      cycle, which H divides); the initial state is slot 0 */
	/* Update order of a model with H > 1 that reads its sources' state at
	 * slot (c - (d-1)) % H through the all-synapse reverse matrix: a synapse
	 * with (d - 1) % H == H - 1 reads slot (c + 1) % H, the slot the step
	 * writes.  NeMo's CPU backend updates the neurons one after another in
	 * index order (cpu/plugins/Kuramoto.cpp:100-126), so such a read sees
	 * the source's new state when its index is below the target's and its
	 * old state otherwise.  level_ptr NULL: no synapse reads that slot.
	 * Otherwise the plugin keeps the slot's old contents (in its scratch)
	 * and updates the neurons in `levels` passes, pass l the group-local
	 * indices level_idx[level_ptr[l] .. level_ptr[l+1]); every lower-index
	 * source such a synapse of a neuron reads is updated in an earlier pass.
	 * level_ptr is host memory (levels + 1 entries, for the launches),
	 * level_idx device memory. */
	unsigned levels;
	const uint32_t* level_ptr;
	const uint32_t* level_idx;
} nemo_hip_neuron_group;

/* The all-synapse reverse matrix (construction/RCM.cpp:76-98), built when a
 * model's .ini asks for [rcm] sources, delays or weights: for each device
 * index n, its incoming synapses k in [ptr[n], ptr[n+1]) in the order NeMo's
 * CPU backend builds them (sources ascending, then each source's synapses in
 * insertion order), source_delay[k] = source device index | (delay-1) << 26,
 * weight[k] the synapse's float weight.  ptr is NULL when no model asks. */
typedef struct nemo_hip_rcm {
	const uint32_t* ptr;
	const uint32_t* source_delay;
	const float* weight;
} nemo_hip_rcm;

/* Bits of the status word a launch may set for the host to report after the
 * step (the results of the step are not those of the reference). */
#define NEMO_HIP_STATUS_SINF_RANGE 1u /* a sine argument outside (-120, 120) */

/* STDP accumulation of step r = s-1 for the neurons of an update block's tile
 * (ConnectivityMatrix::accumulateStdp, ConnectivityMatrix.cpp:224-258;
 * StdpProcess::weightChange, StdpProcess.cpp:68-124).  Skipped when s == 0,
 * r < post_window or r == clock->stdp_flushed.  A neuron n whose history
 * after step r has bit post_window set fired post_window steps before r; for
 * each of its plastic incoming synapses k (rcm_ptr[n] <= k < rcm_ptr[n+1],
 * rcm_sd[k] = source | (delay-1) << 26) the presynaptic firing is the
 * source's history after step r shifted right by the delay (empty for 64);
 * the nearest arrival before and after the postsynaptic spike, within the
 * potentiation and depression masks, selects fn_pre[dt] or fn_post[dt] (a tie
 * gives nothing), and the sum of the two regions is added to rcm_acc[k] with
 * 32-bit wrap-around.  examined[tile] counts the synapses looked at. */
typedef struct nemo_hip_stdp {
	const uint32_t* rcm_ptr;   /* NULL: no STDP accumulation in the update */
	const uint32_t* rcm_sd;
	int32_t* rcm_acc;
	unsigned long long* examined;
	unsigned post_window;      /* P = |postfire| */
	uint64_t potentiation, depression, pre_post, post_pre;
	int32_t fn_pre[64];        /* fixed-point weight change at dt = index+1 */
	int32_t fn_post[64];
} nemo_hip_stdp;

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
 * reads the firing history after step t-1.  The accumulation of up to 64
 * past steps may also run in one batch launch (stdp_batch_kernel) from two
 * snapshots of the firing history in stdp_snap, taken after the last of them
 * and after the step before the first.  Skipped
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
	unsigned long long* block_events; /* events delivered, accumulated: [NEMO_MAX_BIN_BLOCKS] per table
	                                     column, then one per 64 device indices (pushed events) */
	/* STDP in its own launch */
	unsigned stdp_words;             /* bitmap words (64 targets) per STDP block */
	uint64_t* hist;                  /* two planes of hist_pitch words (NEMO_HIST_INDEX) */
	size_t hist_pitch;
	uint64_t* stdp_snap;             /* STDP batches: two history snapshot slots per neuron (2n, 2n+1) */
	nemo_hip_stdp stdp;              /* examined: one counter per STDP block */
} nemo_propagate_args;

/* Propagate blocks an update launch carries after its tiles: the delivery
 * blocks of args (args.step_offset names their step, the next one) and then
 * stdp_blocks STDP blocks (of args' step minus two, the step before the
 * update's).  They need lds_bytes of dynamic LDS and NEMO_COLAUNCH_BLOCK
 * threads, which the launch requests for all its blocks.  Both counts are zero when the launch carries none;
 * only the launch of the first neuron group of a step carries any.
 *
 * push (every group's launch of the step, when set): record_firing delivers
 * the delay-1 events of the step's firing itself, for the next step, into
 * that step's overflow accumulators (acc of the next step's parity; args
 * gives fcm, acc and acc_pitch; push_rows the delay-1 row of every device
 * index, which gather_input reads ahead) and counts them in the block_events
 * entry of their wave (NEMO_MAX_BIN_BLOCKS + device index / 64). */
typedef struct nemo_hip_colaunch {
	unsigned deliver_blocks;
	unsigned stdp_blocks;
	unsigned lds_bytes;
	unsigned push;
	const uint2* push_rows;          /* per device index n: fcm [x, y) of row (n, delay 1) */
	nemo_propagate_args args;
} nemo_hip_colaunch;

/* Per-simulation buffers shared by all groups in a step. */
typedef struct nemo_hip_step_buffers {
	uint64_t step_offset;      /* step = clock->base_step + step_offset */
	const nemo_hip_clock* clock;
	/* Delivery output of step s lives in the buffers of parity s % 2: plane p
	 * of bin_table starts at p * tiles * bin_blocks, of bins at
	 * p * bin_blocks * bin_capacity, of acc at p * 2 * acc_pitch. */
	const uint32_t* bin_table; /* [2][tile][bin_blocks]: count << 16 | offset */
	const uint32_t* bin_range; /* [2][tiles]: for each of the two delivery-block sets, per tile,
	                              first | end << 16 of the blocks holding its events */
	const int2* bins;          /* [2][bin_blocks][bin_capacity] {target % NEMO_TILE, Q weight} */
	unsigned bin_blocks;       /* table columns (blocks of both sets) */
	unsigned bin_capacity;
	unsigned tiles;
	unsigned bin_ranged;       /* nonzero: a tile's events lie in its bin_range columns only, and those
	                              are much fewer than bin_blocks (batches): read just them */
	int64_t* acc;              /* [2][2][acc_pitch] overflow accumulators: E plane, then I plane */
	size_t acc_pitch;
	float* ext_current;        /* per device index, or NULL */
	const uint32_t* fstim;     /* firing-stimulus bit-vector, or NULL */
	uint64_t* hist;            /* 64-bit firing history: two planes of hist_pitch words (NEMO_HIST_INDEX) */
	size_t hist_pitch;
	uint64_t* fired_bits;      /* NEMO_STEP_RING bitmaps of fired_words words */
	size_t fired_words;
	unsigned fbits;            /* fractional bits of the fixed-point currents */
	float* debug_current;      /* NULL, or 2 planes of acc_pitch floats receiving E and I */
	nemo_hip_stdp stdp;
	nemo_hip_colaunch co;      /* extra blocks of this launch (the first group's only) */
	hipEvent_t time_start;     /* when not NULL, recorded when the launch's kernel starts */
	hipEvent_t time_stop;      /* when not NULL, recorded when it ends (hipExtLaunchKernelGGL) */
	nemo_hip_rcm rcm_all;      /* all-synapse reverse matrix, for the models that read it */
	unsigned* status;          /* NEMO_HIP_STATUS_* bits (atomicOr) */
	const float* ext_held;     /* the external current of every device index, also in steps
	                              without a current stimulus (ext_current NULL): for a model
	                              that reads it without clearing it, as a held value */
	unsigned bin_first_set;    /* nonzero: this step's binned events lie in the first delivery-block
	                              set's columns only, [0, bin_first_set) of each table row (with
	                              bin_ranged: in the tile's first range only); the delay-1 events
	                              were pushed into acc by the previous step's update */
} nemo_hip_step_buffers;

typedef hipError_t hip_update_neurons_t(hipStream_t stream,
		const nemo_hip_neuron_group* group, const nemo_hip_step_buffers* buffers);

typedef hipError_t hip_init_neurons_t(hipStream_t stream,
		const nemo_hip_neuron_group* group, const nemo_hip_step_buffers* buffers);

#ifdef __cplusplus
}
#endif

#endif
