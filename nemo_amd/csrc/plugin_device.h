/*
 * Device helpers shared by the neuron-model plugin kernels.
 *
 * An update kernel runs one block of NEMO_TILE threads per tile of its neuron
 * type, one neuron per thread, and every thread of a block calls
 *   UpdateLds& L = update_lds<CO>();            (CO: the launch carries propagate blocks)
 *   h = load_history(b, ng, valid);
 *   gather_input(b, g, L);                      (block barriers inside)
 * or, in a launch of NEMO_COLAUNCH_BLOCK-thread blocks whose tile blocks keep
 * their other waves (colaunched_block<true>), those waves help to gather:
 *   if(!gather_input<NEMO_COLAUNCH_BLOCK>(b, g, L)) return;
 *   take_current(b, L.input, ng, E, I);         (valid threads)
 *   record_firing(b, L, ng, valid, fired, h);   (every thread; pushes the delay-1 events when asked)
 *   stdp_tile(b, L, ng, valid, h);              (every thread; barriers inside)
 */
#ifndef NEMO_AMD_PLUGIN_DEVICE_H
#define NEMO_AMD_PLUGIN_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nemo_hip_plugin.h"
#include "glibc_math.h"

#define NEMO_UPDATE_BLOCK NEMO_TILE
/* Launch bounds of an update kernel, compiled with CO true for launches that
 * carry propagate blocks and false otherwise. */
#define NEMO_UPDATE_BOUNDS(CO) \
	__launch_bounds__((CO) ? NEMO_COLAUNCH_BLOCK : NEMO_UPDATE_BLOCK, (CO) ? NEMO_COLAUNCH_WAVES : 1)

namespace nemo_dev {

__device__ inline uint64_t current_step(const nemo_hip_step_buffers& b)
{
	return b.clock->base_step + b.step_offset;
}

/* Exclusive scan of one value per thread over a block of THREADS threads;
 * returns the thread's prefix and sets *total. */
template<unsigned THREADS = NEMO_TILE>
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* total)
{
	__shared__ uint32_t s_wave[THREADS / 64];
	const unsigned lane = threadIdx.x & 63;
	const unsigned wave = threadIdx.x >> 6;
	uint32_t v = x;
#pragma unroll
	for(int o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(v, o);
		if(int(lane) >= o) v += y;
	}
	if(lane == 63) s_wave[wave] = v;
	__syncthreads();
	uint32_t before = 0, all = 0;
#pragma unroll
	for(unsigned w = 0; w < THREADS / 64; ++w) {
		const uint32_t t = s_wave[w];
		before += w < wave ? t : 0;
		all += t;
	}
	__syncthreads();
	*total = all;
	return before + v - x;
}

/* The largest i < n with arr[i] <= x, for a non-decreasing arr with
 * arr[0] <= x and n <= MAXN (a power of two): a fixed number of unrolled
 * steps, so that searches of independent items interleave. */
template<unsigned MAXN>
__device__ inline unsigned last_not_above(const uint32_t* arr, unsigned n, uint32_t x)
{
	unsigned lo = 0;
#pragma unroll
	for(unsigned step = MAXN / 2; step >= 1; step >>= 1) {
		const unsigned probe = lo + step;
		if(probe < n && arr[probe] <= x) lo = probe;
	}
	return lo;
}

/* last_not_above for K items at once, step by step in lock-step, so that
 * the K LDS reads of a step are in flight together.  Requires n >= 1. */
template<unsigned MAXN, int K>
__device__ inline void last_not_above_k(const uint32_t* arr, unsigned n, const uint32_t (&x)[K], unsigned (&lo)[K])
{
#pragma unroll
	for(int u = 0; u < K; ++u) lo[u] = 0;
#pragma unroll
	for(unsigned step = MAXN / 2; step >= 1; step >>= 1) {
		uint32_t v[K];
#pragma unroll
		for(int u = 0; u < K; ++u) {
			const unsigned probe = lo[u] + step;
			v[u] = arr[probe < n ? probe : n - 1];
		}
#pragma unroll
		for(int u = 0; u < K; ++u) {
			const unsigned probe = lo[u] + step;
			lo[u] = (probe < n && v[u] <= x[u]) ? probe : lo[u];
		}
	}
}

/* Synaptic input of one tile for the current step. */
struct TileInput {
	long long E[NEMO_TILE];
	long long I[NEMO_TILE];
};

/* Gather batches: GATHER_ILP events per thread in flight (6 when the tile's
 * block has NEMO_COLAUNCH_BLOCK threads, all of which gather: 3072 events per
 * batch, and 0.25 us per step faster than 8 at 100k x 1000); an event's column
 * is found from buckets of 2^GATHER_SHIFT events, marked per column when the
 * tile has at most GATHER_MARKED events. */
constexpr int GATHER_ILP = 8;
constexpr int GATHER_ILP_WIDE = 6;
constexpr unsigned GATHER_SHIFT = 3;
constexpr unsigned GATHER_MARKED = 4096;
/* events of the largest batch or marked gather */
constexpr unsigned GATHER_BUCKETS_MAX_EVENTS = GATHER_MARKED > GATHER_ILP_WIDE * NEMO_COLAUNCH_BLOCK
	? (GATHER_MARKED > GATHER_ILP * NEMO_TILE ? GATHER_MARKED : GATHER_ILP * NEMO_TILE)
	: GATHER_ILP_WIDE * NEMO_COLAUNCH_BLOCK;
constexpr unsigned GATHER_BUCKETS = GATHER_BUCKETS_MAX_EVENTS / (1u << GATHER_SHIFT) + 2;

/* The LDS an update block uses: its tile's input sums, the gather's column
 * prefix, bases and buckets, and the per-tile STDP lists.  A kernel takes it
 * from update_lds<DYNAMIC>(): a static array, or -- in launches that also run
 * propagate blocks -- the start of the launch's dynamic LDS, which those
 * blocks use for their own purposes, so that a block of either kind needs
 * only the larger of the two. */
struct UpdateLds {
	TileInput input;
	uint32_t pre[NEMO_MAX_BIN_BLOCKS + 1];
	uint32_t base[NEMO_MAX_BIN_BLOCKS];
	uint32_t bucket[GATHER_BUCKETS];
	int32_t fn[128];
	uint32_t rb[NEMO_TILE];
	uint32_t pos[NEMO_TILE + 1];
	uint2 push_rows[NEMO_TILE]; /* the tile's delay-1 rows, read ahead for push_delay1 */
};

template<bool DYNAMIC>
__device__ inline UpdateLds& update_lds()
{
	if constexpr(DYNAMIC) {
		extern __shared__ uint64_t nemo_update_lds_dyn[];
		return *reinterpret_cast<UpdateLds*>(nemo_update_lds_dyn);
	} else {
		__shared__ UpdateLds lds;
		return lds;
	}
}

/* Sum the tile's input of this step into LDS: the overflow accumulators of
 * its neurons (cleared here, cpu/Simulation.cpp:279-288) and the binned events
 * of every delivery column (integer sums, so the order of the 64-bit LDS
 * atomics does not matter).  The whole table row is read: columns whose
 * window does not hold the tile keep a zero count, and reading them costs
 * less than a dependent load of the tile's column range. */
struct NoPrefetch {
	__device__ void operator()() const {}
};

/* The delay-1 row of device index ng when the launch pushes (co.push), for
 * L.push_rows[threadIdx.x]: gather_input reads it ahead; a kernel that
 * records firing without gathering stores it itself. */
__device__ inline uint2 push_row(const nemo_hip_step_buffers& b, unsigned ng)
{
	return b.co.push ? b.co.push_rows[ng] : make_uint2(0, 0);
}

/* prefetch() is called once the tile's table row and accumulators are
 * requested: the loads it issues (the model's state, say) queue behind them,
 * so that waiting for the table does not wait for them too.
 *
 * THREADS threads of the block take part: the tile's NEMO_TILE (one per
 * neuron, which alone read the table row, the accumulators and, through
 * prefetch, their neuron's state) and, in a block of NEMO_COLAUNCH_BLOCK
 * threads, helpers that only load and sum events, so that a tile of up to
 * 3072 events (GATHER_ILP_WIDE per thread) takes one round of loads.  Returns false for
 * the helpers, which take no further part in the step. */
template<unsigned THREADS = NEMO_TILE, typename Prefetch = NoPrefetch>
__device__ inline bool gather_input(const nemo_hip_step_buffers& b, const nemo_hip_neuron_group& g, UpdateLds& L,
		Prefetch prefetch = Prefetch())
{
	static_assert(THREADS % 64 == 0 && THREADS >= NEMO_TILE, "whole waves beyond the tile");
	TileInput& in = L.input;
	uint32_t* s_pre = L.pre;
	uint32_t* s_base = L.base;
	const unsigned tid = threadIdx.x;
	const bool lead = THREADS == NEMO_TILE || tid < NEMO_TILE; /* uniform per wave */
	const unsigned tile = g.start / NEMO_TILE + blockIdx.x;
	const size_t parity = current_step(b) & 1;
	int64_t* acc = b.acc + parity * 2 * b.acc_pitch;
	const unsigned ng = tile * NEMO_TILE + tid;
	const unsigned BB = b.bin_blocks;
	const uint32_t* row = b.bin_table + (parity * b.tiles + tile) * BB;
	const size_t cap = b.bin_capacity;
	const int2* bins = b.bins + parity * BB * cap;
	/* the columns read: the whole row, or (bin_ranged) the tile's two ranges,
	 * column first0 + i for i < n0, then first1 + i - n0 */
	const unsigned BR = b.bin_first_set ? b.bin_first_set : BB; /* columns holding events */
	unsigned B = BR, first0 = 0, n0 = BR, first1 = 0;
	if(b.bin_ranged) {
		const uint32_t r0 = b.bin_range[tile], r1 = b.bin_range[b.tiles + tile];
		first0 = r0 & 0xffffu;
		n0 = (r0 >> 16) - first0;
		first1 = r1 & 0xffffu;
		B = n0 + (b.bin_first_set ? 0u : (r1 >> 16) - first1);
	}
	constexpr unsigned PER = NEMO_MAX_BIN_BLOCKS / NEMO_TILE;
	uint32_t ent[PER];
	unsigned col[PER];
	/* a helper's i is at least NEMO_TILE * PER = NEMO_MAX_BIN_BLOCKS >= B */
#pragma unroll
	for(unsigned k = 0; k < PER; ++k) {
		const unsigned i = tid * PER + k;
		col[k] = i < n0 ? first0 + i : first1 + (i - n0);
		ent[k] = i < B ? row[col[k]] : 0u;
	}
	if(lead) {
		const int64_t oe = acc[ng];
		const int64_t oi = acc[b.acc_pitch + ng];
		const uint2 prow = push_row(b, ng);
		prefetch();
		in.E[tid] = oe;
		in.I[tid] = oi;
		L.push_rows[tid] = prow;
		if(oe) acc[ng] = 0;
		if(oi) acc[b.acc_pitch + ng] = 0;
	}
	if(B == 0) {
		__syncthreads();
		return lead;
	}
	uint32_t mine = 0;
#pragma unroll
	for(unsigned k = 0; k < PER; ++k) {
		const unsigned i = tid * PER + k;
		if(i < B) s_base[i] = uint32_t(col[k] * cap + (ent[k] & 0xffffu));
		mine += ent[k] >> 16;
	}
	/* batches of ILP events per thread, ILP independent loads in flight per
	 * thread before the LDS atomics.  An event's column lies
	 * between the columns of the bucket bounds around it (every 8th event),
	 * which each column marks when the tile has at most MARKED events, or
	 * which are searched per batch; it is refined between them. */
	constexpr int ILP = THREADS > NEMO_TILE ? GATHER_ILP_WIDE : GATHER_ILP;
	constexpr unsigned BATCH = ILP * THREADS;
	constexpr unsigned SHIFT = GATHER_SHIFT;
	constexpr unsigned MARKED = GATHER_MARKED;
	static_assert(BATCH <= GATHER_BUCKETS_MAX_EVENTS, "the bucket array holds a batch");
	uint32_t* s_bucket = L.bucket;
	uint32_t total;
	uint32_t pre = block_exclusive_scan<THREADS>(mine, &total);
	const bool marked = total <= MARKED;
#pragma unroll
	for(unsigned k = 0; k < PER; ++k) {
		const unsigned i = tid * PER + k;
		const uint32_t c = ent[k] >> 16;
		if(i < B) {
			s_pre[i] = pre;
			if(marked) {
				for(uint32_t j = (pre + (1u << SHIFT) - 1) >> SHIFT; (j << SHIFT) < pre + c; ++j) s_bucket[j] = i;
			}
		}
		pre += c;
	}
	if(tid == 0) {
		s_pre[B] = total;
		if(marked && total) s_bucket[((total - 1) >> SHIFT) + 1] = B - 1;
	}
	__syncthreads();
	for(uint32_t base = 0; base < total; base += BATCH) {
		const uint32_t b0 = marked ? 0u : base; /* first event of the buckets */
		if(!marked) {
			for(unsigned j = tid; j <= (BATCH >> SHIFT); j += THREADS) {
				s_bucket[j] = last_not_above<NEMO_MAX_BIN_BLOCKS>(s_pre, B, min(base + (j << SHIFT), total - 1));
			}
			__syncthreads();
		}
		uint32_t jj[ILP];
		unsigned lo[ILP], hi[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			/* buckets past the last event may be unset: only the events of
			 * the tile read theirs */
			jj[u] = base + u * THREADS + tid;
			const unsigned j = (jj[u] - b0) >> SHIFT;
			lo[u] = jj[u] < total ? s_bucket[j] : 0u;
			hi[u] = jj[u] < total ? s_bucket[j + 1] : 0u;
		}
		bool open = true;
		while(open) {
			open = false;
#pragma unroll
			for(int u = 0; u < ILP; ++u) {
				if(lo[u] < hi[u]) {
					const unsigned mid = lo[u] + ((hi[u] - lo[u] + 1) >> 1);
					if(s_pre[mid] <= jj[u]) lo[u] = mid; else hi[u] = mid - 1;
					open = open || lo[u] < hi[u];
				}
			}
		}
		int2 ev[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			ev[u] = make_int2(0, 0);
			if(jj[u] < total) ev[u] = bins[s_base[lo[u]] + (jj[u] - s_pre[lo[u]])];
		}
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			if(jj[u] < total) {
				long long* dst = ev[u].y >= 0 ? &in.E[ev[u].x] : &in.I[ev[u].x];
				atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)(long long)ev[u].y);
			}
		}
		if(!marked) __syncthreads(); /* the buckets are rewritten by the next batch */
	}
	__syncthreads();
	return lead;
}

/* Saturate a 64-bit fixed-point sum to the int32 range and convert to float,
 * as NeMo's wfx_toFloat (src/nemo/fixedpoint.cpp:40-49). */
__device__ inline float wfx_to_float(int64_t v, unsigned fbits)
{
	v = v > int64_t(INT32_MAX) ? int64_t(INT32_MAX) : v;
	v = v < int64_t(INT32_MIN) ? int64_t(INT32_MIN) : v;
	return float(int32_t(v)) / float(1u << fbits);
}

/* Excitatory and inhibitory current of neuron ng (thread threadIdx.x of the
 * tile), as gathered by gather_input. */
__device__ inline void take_current(const nemo_hip_step_buffers& b, const TileInput& in, unsigned ng, float& E, float& I)
{
	E = wfx_to_float(in.E[threadIdx.x], b.fbits);
	I = wfx_to_float(in.I[threadIdx.x], b.fbits);
	if(b.debug_current) {
		b.debug_current[ng] = E;
		b.debug_current[b.acc_pitch + ng] = I;
	}
}

/* User current for this step, cleared after use (Izhikevich.cpp:61-64). */
__device__ inline float take_external(const nemo_hip_step_buffers& b, unsigned ng)
{
	if(!b.ext_current) return 0.0f;
	float v = b.ext_current[ng];
	b.ext_current[ng] = 0.0f;
	return v;
}

__device__ inline bool stimulated(const nemo_hip_step_buffers& b, unsigned ng)
{
	return b.fstim && ((b.fstim[ng >> 5] >> (ng & 31)) & 1u);
}

/* Marsaglia xorshift128 (src/nemo/RNG.cpp:9-19) */
__device__ inline uint32_t urand(uint4& s)
{
	uint32_t t = s.x ^ (s.x << 11);
	s.x = s.y;
	s.y = s.z;
	s.z = s.w;
	s.w = (s.w ^ (s.w >> 19)) ^ (t ^ (t >> 8));
	return s.w;
}

/* Box-Muller, sine branch only, bit-exact with glibc (src/nemo/RNG.cpp:25-33) */
__device__ inline float nrand(uint4& s)
{
	float a = float(urand(s)) * 1.4629180792671596810513378043098e-9f;
	float b = float(urand(s)) * 0.00000000023283064365386962890625f;
	float r = __builtin_sqrtf(-2.0f * nemo_math::logf_exact(1.0f - b));
	return nemo_math::sinf_exact(a) * r;
}

/* Write the tile's words of the step's firing bitmap (one ballot per wave;
 * device indices of a wave are 64-aligned). */
__device__ inline void store_firing(const nemo_hip_step_buffers& b, unsigned ng, bool valid, bool fired)
{
	const uint64_t mask = __ballot(valid && fired);
	if((threadIdx.x & 63) == 0) {
		const uint64_t slot = current_step(b) & (NEMO_STEP_RING - 1);
		b.fired_bits[slot * b.fired_words + (ng >> 6)] = mask;
	}
}

/* History of neuron ng before this step (after the previous one). */
__device__ inline uint64_t load_history(const nemo_hip_step_buffers& b, unsigned ng, bool valid)
{
	return valid ? b.hist[NEMO_HIST_INDEX(ng, current_step(b) - 1, b.hist_pitch)] : 0;
}

/* Delay-1 delivery of this step's firing for the next step, straight into
 * that step's overflow accumulators (nemo_hip_colaunch.push), as the delay-1
 * part of cpu::Simulation::deliverSpikes (src/nemo/cpu/Simulation.cpp:
 * 259-308).  Each wave spreads the delay-1 rows of its fired neurons (read
 * ahead into L.push_rows by gather_input) over its lanes, PUSH_ILP loads in
 * flight per lane before their 64-bit atomics; the sums are integer, so
 * their order does not matter.  Every thread of the wave calls it. */
constexpr int PUSH_ILP = 4;
__device__ inline void push_delay1(const nemo_hip_step_buffers& b, const UpdateLds& L, unsigned ng, bool valid,
		bool fired)
{
	const nemo_propagate_args& a = b.co.args;
	const bool f = valid && fired;
	if(!__any(f)) return;
	const unsigned lane = threadIdx.x & 63;
	uint32_t rb = 0, len = 0;
	if(f) {
		const uint2 r = L.push_rows[threadIdx.x];
		rb = r.x;
		len = r.y - r.x;
	}
	uint32_t incl = len;
#pragma unroll
	for(int o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(incl, o);
		if(int(lane) >= o) incl += y;
	}
	const uint32_t total = __shfl(incl, 63);
	const uint32_t excl = incl - len;
	int64_t* acc = a.acc + ((current_step(b) + 1) & 1) * 2 * a.acc_pitch;
	for(uint32_t e0 = 0; e0 < total; e0 += 64 * PUSH_ILP) {
		int2 ev[PUSH_ILP];
#pragma unroll
		for(int u = 0; u < PUSH_ILP; ++u) {
			const uint32_t e = e0 + u * 64 + lane;
			/* the lane whose row holds event e: the first whose inclusive
			 * prefix exceeds e (the prefix is non-decreasing over lanes) */
			unsigned s = 0;
#pragma unroll
			for(unsigned step = 32; step >= 1; step >>= 1) {
				if(__shfl(incl, int(s + step - 1)) <= e) s += step;
			}
			const uint32_t srb = __shfl(rb, int(s));
			const uint32_t sex = __shfl(excl, int(s));
			ev[u] = make_int2(0, 0);
			if(e < total) ev[u] = a.fcm[srb + (e - sex)];
		}
#pragma unroll
		for(int u = 0; u < PUSH_ILP; ++u) {
			if(e0 + u * 64 + lane < total) {
				int64_t* dst = acc + (ev[u].y >= 0 ? size_t(0) : a.acc_pitch) + uint32_t(ev[u].x);
				atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)(int64_t)ev[u].y);
			}
		}
	}
	/* the wave's own counter: atomics of many waves on one address queue */
	if(lane == 0) atomicAdd(&a.block_events[NEMO_MAX_BIN_BLOCKS + (ng >> 6)], (unsigned long long)total);
}

/* Store the shifted history (h as from load_history) and write the step's
 * firing bitmap. */
__device__ inline void record_firing(const nemo_hip_step_buffers& b, unsigned ng, bool valid, bool fired, uint64_t h)
{
	if(valid) {
		b.hist[NEMO_HIST_INDEX(ng, current_step(b), b.hist_pitch)] = (h << 1) | uint64_t(fired ? 1 : 0);
	}
	store_firing(b, ng, valid, fired);
}

/* The same in an update block whose input gather_input summed into L: with
 * co.push, also deliver the firing's delay-1 events. */
__device__ inline void record_firing(const nemo_hip_step_buffers& b, const UpdateLds& L, unsigned ng, bool valid,
		bool fired, uint64_t h)
{
	record_firing(b, ng, valid, fired, h);
	if(b.co.push) push_delay1(b, L, ng, valid, fired);
}

/* StdpProcess::updateRegion (StdpProcess.cpp:86-111): within one region of
 * arrivals, the nearest presynaptic arrival on either side of the
 * postsynaptic spike wins; a tie gives nothing. */
__device__ inline int32_t stdp_update_region(uint64_t arrivals, unsigned post_window, uint64_t pre_post,
		uint64_t post_pre, const int32_t* fn_pre, const int32_t* fn_post)
{
	if(!arrivals) return 0;
	const unsigned NONE = ~0u;
	const uint64_t pre = arrivals & pre_post;
	const uint64_t post = arrivals & post_pre;
	const unsigned dt_pre = pre ? unsigned(__builtin_ctzll(pre >> post_window)) : NONE;
	const unsigned dt_post = post ? unsigned(__builtin_clzll(post << (64 - post_window))) : NONE;
	if(dt_pre < dt_post) return fn_pre[dt_pre];
	if(dt_post < dt_pre) return fn_post[dt_post];
	return 0;
}

__device__ inline int32_t stdp_wrap_add(int32_t x, int32_t y)
{
	return int32_t(uint32_t(x) + uint32_t(y));
}

/* The STDP accumulation of the previous step for this tile's neurons (the
 * contract of nemo_hip_stdp).  h is the thread's neuron history after the
 * previous step (load_history).  Every thread of the block calls it. */
__device__ inline void stdp_tile(const nemo_hip_step_buffers& b, UpdateLds& L, unsigned ng, bool valid, uint64_t h)
{
	constexpr int ILP = 8;
	const nemo_hip_stdp& st = b.stdp;
	if(!st.rcm_ptr) return;
	const uint64_t t = current_step(b);
	if(t == 0) return;
	const uint64_t r = t - 1;
	if(r < st.post_window || r == b.clock->stdp_flushed) return;

	int32_t* s_fn = L.fn;
	uint32_t* s_rb = L.rb;
	uint32_t* s_pos = L.pos;
	const unsigned tid = threadIdx.x;
	if(tid < 64) {
		s_fn[tid] = st.fn_pre[tid];
		s_fn[64 + tid] = st.fn_post[tid];
	}
	const bool post = valid && ((h >> st.post_window) & 1u);
	uint32_t rb = 0, len = 0;
	if(post) {
		rb = st.rcm_ptr[ng];
		len = st.rcm_ptr[ng + 1] - rb;
	}
	uint32_t nfired;
	const uint32_t idx = block_exclusive_scan(post ? 1u : 0u, &nfired);
	uint32_t total;
	const uint32_t pos = block_exclusive_scan(len, &total);
	if(post) {
		s_rb[idx] = rb;
		s_pos[idx] = pos;
	}
	if(tid == 0) s_pos[nfired] = total;
	__syncthreads();
	if(total == 0) return;

	const uint64_t* pre_hist = b.hist + NEMO_HIST_INDEX(0, r, b.hist_pitch);
	for(uint32_t e0 = tid; e0 < total; e0 += ILP * NEMO_TILE) {
		uint32_t k[ILP];
		uint32_t sd[ILP];
		int32_t acc[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) k[u] = e0 + u * NEMO_TILE;
		unsigned lo[ILP];
		last_not_above_k<NEMO_TILE>(s_pos, nfired, k, lo);
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			const uint32_t e = k[u];
			k[u] = 0;
			sd[u] = 0;
			acc[u] = 0;
			if(e < total) {
				k[u] = s_rb[lo[u]] + (e - s_pos[lo[u]]);
				sd[u] = st.rcm_sd[k[u]];
				acc[u] = st.rcm_acc[k[u]];
			}
		}
		uint64_t ph[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			ph[u] = 0;
			if(e0 + u * NEMO_TILE < total) ph[u] = pre_hist[sd[u] & 0x3ffffffu];
		}
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			if(e0 + u * NEMO_TILE < total) {
				const unsigned d = (sd[u] >> 26) + 1;
				const uint64_t arrivals = d >= 64 ? 0 : (ph[u] >> d);
				const int32_t dw = stdp_wrap_add(
						stdp_update_region(arrivals & st.potentiation, st.post_window, st.pre_post, st.post_pre,
							s_fn, s_fn + 64),
						stdp_update_region(arrivals & st.depression, st.post_window, st.pre_post, st.post_pre,
							s_fn, s_fn + 64));
				if(dw != 0) st.rcm_acc[k[u]] = stdp_wrap_add(acc[u], dw);
			}
		}
	}
	if(tid == 0) atomicAdd(&st.examined[ng / NEMO_TILE], (unsigned long long)total);
}

} // namespace nemo_dev

/* the propagate blocks an update launch may carry (colaunched_block) */
#include "propagate_device.h"

#endif
