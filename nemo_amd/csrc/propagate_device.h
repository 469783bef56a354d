/*
 * Device code of the propagate blocks -- spike delivery of one step and STDP
 * accumulation of the previous one -- shared by the backend's own propagate
 * launches and by neuron-update launches that carry them as extra blocks
 * after their tiles.  Templated on the block width; all LDS comes from one
 * dynamic buffer whose layout PropagateLds describes, so a launch that also
 * runs other blocks pays no static LDS for these.
 *
 * Semantics: delivery as cpu::Simulation::deliverSpikes
 * (src/nemo/cpu/Simulation.cpp:259-308), binned by target tile (the contract
 * of nemo_propagate_args); STDP as ConnectivityMatrix::accumulateStdp
 * (src/nemo/ConnectivityMatrix.cpp:224-258) with StdpProcess::weightChange
 * (src/nemo/StdpProcess.cpp:68-124).
 */
#ifndef NEMO_AMD_PROPAGATE_DEVICE_H
#define NEMO_AMD_PROPAGATE_DEVICE_H

#include "plugin_device.h"

namespace nemo_dev {

constexpr uint64_t PROPAGATE_RING_MASK = NEMO_STEP_RING - 1;
/* independent global loads each thread keeps in flight in the synapse loop
 * of STDP, and in the event loop of delivery */
constexpr int ILP = 8;
/* 8, not 16: with the launch held to 80 VGPRs a CU runs three 512-thread
 * delivery blocks at once instead of two, which pays when there are more
 * blocks than CUs (batches) and costs nothing with one block per CU */
constexpr int EVENT_ILP = 8;
/* The delivery blocks an update launch carries (overlap 3) keep 6: that
 * launch is held to 72 VGPRs (NEMO_COLAUNCH_WAVES), so that beside three of
 * its blocks a CU's SIMDs still hold one wave each of an STDP batch block
 * (NEMO_STDP_BATCH_BLOCK threads, 80 VGPRs); with 8 the launch needs 78, and
 * a batch block beside it leaves a CU room for one of its blocks instead of
 * three, so that the blocks of the steps beside a batch wait for each other */
constexpr int CARRIED_EVENT_ILP = 6;
/* independent global loads per thread in the synapse loop of the STDP batch
 * beside the steps, held to 80 VGPRs (NEMO_STDP_BATCH_WAVES waves per SIMD),
 * and of a flush's batch, which nothing runs beside */
constexpr int STDP_BATCH_ILP = 5;
constexpr int STDP_FLUSH_ILP = 8;
/* staged events per thread the delivery sort keeps in registers */
constexpr unsigned SORT_REGS = 8;

/* LDS of a propagate block of THREADS threads, carved from one buffer:
 * staged events (bin capacity), per-window-tile counts and cursors, the
 * bitmap words of a batch with their prefix, and the fired rows of a chunk
 * with their event prefix.  An STDP block uses the start of the same buffer
 * for its weight-change table, row bases and prefix. */
template<unsigned THREADS, int EV_ILP = EVENT_ILP>
struct PropagateLds {
	/* fired rows a delivery block lists at a time */
	static constexpr unsigned ROWS = THREADS < 512 ? 512 : THREADS;
	int2* stage;
	uint64_t* word;
	uint32_t* wpre;
	uint32_t* rb;
	uint32_t* rpre;
	uint32_t* cnt;
	uint32_t* cur;
	uint32_t* bucket;
	/* events of a batch of the event loop, events per row bucket, and the
	 * most events of a chunk whose buckets the rows mark (buckets hold the
	 * larger of a batch and that) */
	static constexpr unsigned BATCH = EV_ILP * THREADS;
	static constexpr unsigned BUCKET_SHIFT = 5;
	static constexpr unsigned MARKED = 16384;
	static constexpr unsigned BUCKETS = (BATCH > MARKED ? BATCH : MARKED) >> BUCKET_SHIFT;
	__device__ PropagateLds(unsigned char* p, unsigned cap, unsigned max_window)
	{
		stage = reinterpret_cast<int2*>(p);
		word = reinterpret_cast<uint64_t*>(stage + cap);
		wpre = reinterpret_cast<uint32_t*>(word + THREADS);
		rb = wpre + THREADS + 2;
		rpre = rb + ROWS;
		bucket = rpre + ROWS + 2;
		cnt = bucket + BUCKETS + 2;
		cur = cnt + max_window;
	}
	static constexpr size_t delivery_bytes(unsigned cap, unsigned max_window)
	{
		return size_t(cap) * 8 + size_t(THREADS) * 8 + 4 * (size_t(THREADS) + 2) + 4 * (2 * size_t(ROWS) + 2)
			+ 4 * (size_t(BUCKETS) + 2) + 8 * size_t(max_window);
	}
	static constexpr size_t stdp_bytes() { return 128 * 4 + 4 * (2 * size_t(THREADS) + 2) + 8 * size_t(THREADS); }
	static constexpr size_t bytes(bool delivery, unsigned cap, unsigned max_window)
	{
		return delivery && delivery_bytes(cap, max_window) > stdp_bytes() ? delivery_bytes(cap, max_window)
			: stdp_bytes();
	}
};

__device__ inline unsigned kth_set_bit(uint64_t w, unsigned k)
{
	for(unsigned i = 0; i < k; ++i) w &= w - 1;
	return unsigned(__builtin_ctzll(w));
}

/* ------------------------------------------------------------ delivery */

/* A delivery block: list the fired rows of its (source tile, delay) pairs,
 * read their row pointers in parallel, then spread all their events evenly
 * over the block (four loads in flight per thread), staging them in LDS;
 * finally counting-sort the staged events by target tile (relative to the
 * block's window) into the block's bins. */
template<unsigned THREADS, int EV_ILP = EVENT_ILP>
__device__ inline void deliver_block(const nemo_propagate_args& a, uint64_t t, unsigned b, unsigned char* lds)
{
	using Lds = PropagateLds<THREADS, EV_ILP>;
	constexpr unsigned ROWS = Lds::ROWS;
	constexpr unsigned ROWS_PER_THREAD = ROWS / THREADS;
	const Lds L(lds, a.bin_capacity, a.max_window);
	int2* stage = L.stage;
	uint32_t* cnt = L.cnt;
	uint32_t* cur = L.cur;
	uint64_t* s_word = L.word;
	uint32_t* s_wpre = L.wpre;
	uint32_t* s_rb = L.rb;
	uint32_t* s_rpre = L.rpre;

	const unsigned CAP = a.bin_capacity;
	const unsigned D = a.D;
	const unsigned DC = a.d_count;
	const unsigned tid = threadIdx.x;
	const unsigned col = a.block_base + b;
	const size_t parity = t & 1;
	int64_t* acc = a.acc + parity * 2 * a.acc_pitch;
	const uint2 win = a.windows[b];
	const unsigned tlo = win.x;
	const unsigned W = win.y - win.x; /* tiles in the window */
	for(unsigned p = tid; p < W; p += THREADS) cnt[p] = 0;

	const uint64_t q0 = a.pair_bounds[b];
	const uint64_t q1 = a.pair_bounds[b + 1];
	const uint64_t nwords = (q1 - q0) * 4;
	uint32_t fill = 0;
	unsigned long long events = 0;

	for(uint64_t wb = 0; wb < nwords; wb += THREADS) {
		/* the bitmap words of this batch, one per thread */
		uint64_t word = 0;
		const uint64_t w = wb + tid;
		if(w < nwords) {
			const uint64_t q = q0 + w / 4;
			const unsigned p = unsigned(q / DC);
			const unsigned d = unsigned(q % DC) + a.d_first;
			if(t >= d) word = a.fired_bits[((t - d) & PROPAGATE_RING_MASK) * a.fired_words + size_t(p) * 4 + (w & 3)];
		}
		__syncthreads(); /* previous batch done with s_word */
		s_word[tid] = word;
		uint32_t R;
		const uint32_t wpre = block_exclusive_scan<THREADS>(uint32_t(__popcll(word)), &R);
		s_wpre[tid] = wpre;
		if(tid == 0) s_wpre[THREADS] = R;
		__syncthreads();

		for(uint32_t r0 = 0; r0 < R; r0 += ROWS) {
			const uint32_t nr = min(ROWS, R - r0);
			/* row pointers of up to ROWS fired rows, ROWS_PER_THREAD per thread */
			uint32_t lens[ROWS_PER_THREAD];
			uint32_t mine = 0;
			uint32_t rr[ROWS_PER_THREAD];
			unsigned wl[ROWS_PER_THREAD];
#pragma unroll
			for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) rr[k] = r0 + tid * ROWS_PER_THREAD + k;
			last_not_above_k<THREADS>(s_wpre, THREADS, rr, wl);
			uint32_t rbv[ROWS_PER_THREAD], rev[ROWS_PER_THREAD];
#pragma unroll
			for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) {
				const uint32_t ri = tid * ROWS_PER_THREAD + k;
				rbv[k] = 0;
				rev[k] = 0;
				if(ri < nr) {
					const unsigned lo = wl[k];
					const uint64_t wg = wb + lo;
					const uint64_t q = q0 + wg / 4;
					const unsigned p = unsigned(q / DC);
					const unsigned d = unsigned(q % DC) + a.d_first;
					const unsigned bit = kth_set_bit(s_word[lo], rr[k] - s_wpre[lo]);
					const uint32_t n = p * NEMO_TILE + unsigned(wg & 3) * 64 + bit;
					const size_t row = size_t(n) * D + (d - 1);
					rbv[k] = a.rowptr[row];
					rev[k] = a.rowptr[row + 1];
				}
			}
#pragma unroll
			for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) {
				const uint32_t ri = tid * ROWS_PER_THREAD + k;
				lens[k] = rev[k] - rbv[k];
				if(ri < nr) s_rb[ri] = rbv[k];
				mine += lens[k];
			}
			uint32_t M;
			uint32_t pre = block_exclusive_scan<THREADS>(mine, &M);
#pragma unroll
			for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) {
				const uint32_t ri = tid * ROWS_PER_THREAD + k;
				if(ri < nr) s_rpre[ri] = pre;
				pre += lens[k];
			}
			if(tid == 0) s_rpre[nr] = M;
			__syncthreads();

			/* staging of event number e of this chunk: into the stage with a
			 * count for its tile, or (beyond capacity or, defensively,
			 * outside the window) straight into the accumulators */
			auto stage_event = [&](int2 ev, uint32_t e) {
				const uint32_t pos = fill + e;
				const uint32_t rel = uint32_t(ev.x) / NEMO_TILE - tlo;
				if(pos < CAP && rel < W) {
					stage[pos] = ev;
					atomicAdd(&cnt[rel], 1u);
				} else {
					if(pos < CAP) stage[pos] = make_int2(-1, 0);
					int64_t* dst = acc + (ev.y >= 0 ? size_t(0) : a.acc_pitch) + uint32_t(ev.x);
					atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)(int64_t)ev.y);
				}
			};
			/* the events of these rows evenly over the block, in batches of
			 * EV_ILP per thread.  An event's row lies between the rows of
			 * the bucket bounds around it (every 32nd event), which each row
			 * marks when the chunk has at most MARKED events, or which are
			 * searched per batch; it is refined between them. */
			constexpr unsigned BATCH = Lds::BATCH;
			constexpr unsigned SHIFT = Lds::BUCKET_SHIFT;
			const bool marked = M <= Lds::MARKED;
			if(marked) {
#pragma unroll
				for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) {
					const uint32_t ri = tid * ROWS_PER_THREAD + k;
					if(ri < nr) {
						const uint32_t p0 = s_rpre[ri], p1 = s_rpre[ri + 1];
						for(uint32_t j = (p0 + (1u << SHIFT) - 1) >> SHIFT; (j << SHIFT) < p1; ++j) L.bucket[j] = ri;
					}
				}
				if(tid == 0 && M) L.bucket[((M - 1) >> SHIFT) + 1] = nr - 1;
				__syncthreads();
			}
			for(uint32_t base = 0; base < M; base += BATCH) {
				const uint32_t b0 = marked ? 0u : base; /* first event of the buckets */
				if(!marked) {
					for(unsigned j = tid; j <= (BATCH >> SHIFT); j += THREADS) {
						L.bucket[j] = last_not_above<ROWS>(s_rpre, nr, min(base + (j << SHIFT), M - 1));
					}
					__syncthreads();
				}
				uint32_t ee[EV_ILP];
				unsigned lo[EV_ILP], hi[EV_ILP];
#pragma unroll
				for(int u = 0; u < EV_ILP; ++u) {
					/* buckets past the last event may be unset: only the
					 * events of the chunk read theirs */
					ee[u] = base + u * THREADS + tid;
					const unsigned j = (ee[u] - b0) >> SHIFT;
					lo[u] = ee[u] < M ? L.bucket[j] : 0u;
					hi[u] = ee[u] < M ? L.bucket[j + 1] : 0u;
				}
				bool open = true;
				while(open) {
					open = false;
#pragma unroll
					for(int u = 0; u < EV_ILP; ++u) {
						if(lo[u] < hi[u]) {
							const unsigned mid = lo[u] + ((hi[u] - lo[u] + 1) >> 1);
							if(s_rpre[mid] <= ee[u]) lo[u] = mid; else hi[u] = mid - 1;
							open = open || lo[u] < hi[u];
						}
					}
				}
				int2 ev[EV_ILP];
#pragma unroll
				for(int u = 0; u < EV_ILP; ++u) {
					ev[u] = make_int2(0, 0);
					if(ee[u] < M) ev[u] = a.fcm[s_rb[lo[u]] + (ee[u] - s_rpre[lo[u]])];
				}
#pragma unroll
				for(int u = 0; u < EV_ILP; ++u) {
					if(ee[u] < M) stage_event(ev[u], ee[u]);
				}
				if(!marked) __syncthreads(); /* the buckets are rewritten by the next batch */
			}
			fill += M;
			events += M;
			__syncthreads(); /* s_rb / s_rpre are rewritten by the next chunk */
		}
	}

	/* counting sort of the staged events by target tile */
	const uint32_t staged = min(fill, CAP);
	uint32_t carry = 0;
	for(uint32_t base = 0; base < W; base += THREADS) {
		const uint32_t p = base + tid;
		const uint32_t c = p < W ? cnt[p] : 0u;
		uint32_t total;
		const uint32_t pre = block_exclusive_scan<THREADS>(c, &total);
		if(p < W) {
			cur[p] = carry + pre;
			a.bin_table[(parity * a.tiles + tlo + p) * a.table_columns + col] = (c << 16) | (carry + pre);
		}
		carry += total;
	}
	__syncthreads();
	int2* out = a.bins + (parity * a.table_columns + col) * CAP;
	if(staged <= SORT_REGS * THREADS) {
		/* every thread holds its staged events in registers, then all write
		 * them back sorted into the stage, which is copied out contiguously */
		int2 e[SORT_REGS];
#pragma unroll
		for(unsigned u = 0; u < SORT_REGS; ++u) {
			const uint32_t i = u * THREADS + tid;
			e[u] = i < staged ? stage[i] : make_int2(-1, 0);
		}
		uint32_t pos[SORT_REGS];
#pragma unroll
		for(unsigned u = 0; u < SORT_REGS; ++u) {
			if(e[u].x >= 0) pos[u] = atomicAdd(&cur[uint32_t(e[u].x) / NEMO_TILE - tlo], 1u);
		}
		__syncthreads();
#pragma unroll
		for(unsigned u = 0; u < SORT_REGS; ++u) {
			if(e[u].x >= 0) stage[pos[u]] = make_int2(int(uint32_t(e[u].x) % NEMO_TILE), e[u].y);
		}
		__syncthreads();
		/* two events per 16-byte store (the stage and every bin region start
		 * 16-byte aligned), the odd last one alone; non-temporal, since the
		 * update reads them on other XCDs */
		typedef int v4i __attribute__((ext_vector_type(4)));
		const uint32_t pairs = carry / 2;
		const v4i* s4 = reinterpret_cast<const v4i*>(stage);
		v4i* o4 = reinterpret_cast<v4i*>(out);
		for(uint32_t i = tid; i < pairs; i += THREADS) __builtin_nontemporal_store(s4[i], &o4[i]);
		if((carry & 1) && tid == 0) out[carry - 1] = stage[carry - 1];
	} else {
		/* four staged events per thread at a time, so that their LDS
		 * atomics (which return the slot) overlap */
		for(uint32_t i0 = tid; i0 < staged; i0 += 4 * THREADS) {
			int2 e[4];
			uint32_t pos[4];
#pragma unroll
			for(int u = 0; u < 4; ++u) {
				const uint32_t i = i0 + u * THREADS;
				e[u] = i < staged ? stage[i] : make_int2(-1, 0);
			}
#pragma unroll
			for(int u = 0; u < 4; ++u) {
				if(e[u].x >= 0) pos[u] = atomicAdd(&cur[uint32_t(e[u].x) / NEMO_TILE - tlo], 1u);
			}
#pragma unroll
			for(int u = 0; u < 4; ++u) {
				if(e[u].x >= 0) out[pos[u]] = make_int2(int(uint32_t(e[u].x) % NEMO_TILE), e[u].y);
			}
		}
	}
	if(tid == 0 && events) a.block_events[col] += events;
}

/* ---------------------------------------------------------------- STDP */

/* An STDP block of the STDP-only launch: the targets among its bitmap words
 * that fired P steps before step r, and all their plastic incoming synapses
 * spread evenly over its threads (ILP loads in flight per thread).  The
 * update kernels do the same per tile when STDP runs inside them. */
template<unsigned THREADS>
__device__ inline void stdp_block(const nemo_propagate_args& a, uint64_t r, unsigned sb, unsigned char* lds)
{
	const nemo_hip_stdp& st = a.stdp;
	int32_t* s_fn = reinterpret_cast<int32_t*>(lds);
	uint32_t* s_rb = reinterpret_cast<uint32_t*>(lds + 128 * sizeof(int32_t));
	uint32_t* s_pos = s_rb + THREADS;
	const unsigned P = st.post_window;
	const unsigned tid = threadIdx.x;
	if(tid < 64) {
		s_fn[tid] = st.fn_pre[tid];
		s_fn[64 + tid] = st.fn_post[tid];
	}

	const uint64_t slot = (r - P) & PROPAGATE_RING_MASK;
	const size_t w0 = size_t(sb) * a.stdp_words;
	const size_t word = w0 + tid / 64;
	bool fired = false;
	if(tid / 64 < a.stdp_words && word < a.fired_words) {
		fired = (a.fired_bits[slot * a.fired_words + word] >> (tid & 63)) & 1u;
	}
	uint32_t nfired;
	const uint32_t idx = block_exclusive_scan<THREADS>(fired ? 1u : 0u, &nfired);
	uint32_t len = 0, rb = 0;
	if(fired) {
		const uint32_t n = uint32_t(w0 * 64 + tid);
		rb = st.rcm_ptr[n];
		len = st.rcm_ptr[n + 1] - rb;
	}
	uint32_t total;
	const uint32_t pos = block_exclusive_scan<THREADS>(len, &total);
	if(fired) {
		s_rb[idx] = rb;
		s_pos[idx] = pos;
	}
	if(tid == 0) s_pos[nfired] = total;
	__syncthreads();
	if(total == 0) return;

	const uint64_t* hist = a.hist + NEMO_HIST_INDEX(0, r, a.hist_pitch);
	for(uint32_t e0 = tid; e0 < total; e0 += ILP * THREADS) {
		uint32_t k[ILP];
		uint32_t sd[ILP];
		int32_t acc[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) k[u] = e0 + u * THREADS;
		unsigned lo[ILP];
		last_not_above_k<THREADS>(s_pos, nfired, k, lo);
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
		uint64_t h[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			h[u] = 0;
			if(e0 + u * THREADS < total) h[u] = hist[sd[u] & 0x3ffffffu];
		}
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			if(e0 + u * THREADS < total) {
				const unsigned d = (sd[u] >> 26) + 1;
				const uint64_t preFiring = d >= 64 ? 0 : (h[u] >> d);
				const int32_t dw = stdp_wrap_add(
						stdp_update_region(preFiring & st.potentiation, P, st.pre_post, st.post_pre,
							s_fn, s_fn + 64),
						stdp_update_region(preFiring & st.depression, P, st.pre_post, st.post_pre,
							s_fn, s_fn + 64));
				if(dw != 0) {
					st.rcm_acc[k[u]] = stdp_wrap_add(acc[u], dw);
				}
			}
		}
	}
	if(tid == 0) atomicAdd(&st.examined[sb], (unsigned long long)total);
}

/* Bits [y, y+64) of the 128-bit firing record c | s << n (1 <= n <= 64,
 * y <= 64), where c is a neuron's history after step rl and s its history
 * after step rl - n: bit x of the record is firing at step rl - x, so this is
 * the history after step rl - y. */
__device__ inline uint64_t history_window(uint64_t c, uint64_t s, unsigned n, unsigned y)
{
	const uint64_t lo = c | (n < 64 ? s << n : 0);
	const uint64_t hi = n < 64 ? s >> (64 - n) : s;
	return y == 0 ? lo : y < 64 ? (lo >> y) | (hi << (64 - y)) : hi;
}

/* The pending STDP accumulation (clock->stdp_first, stdp_count: 1..64
 * steps r0 .. rl) for the targets of block sb's bitmap words, from the two
 * history snapshots only: after rl (slot stdp_snaps % 2 of stdp_snap) and
 * after r0 - 1 (the other slot).  Together they hold every neuron's firing
 * over the 128 steps the n accumulations look at: a target fired at step
 * r - P when bit rl - r + P of its record is set, and a source's history
 * after step r is the window of its record at rl - r.  Each thread owns a
 * synapse for all the steps, so the accumulator is read and written once per
 * batch and no two threads touch it; the sum of 32-bit wrap-around additions
 * does not depend on their order, so the result equals n single-step
 * accumulations.  Nothing here reads the live history, so the batch may run
 * beside the steps that follow. */
template<unsigned THREADS, int SYN_ILP = STDP_BATCH_ILP>
__device__ inline void stdp_batch_block(const nemo_propagate_args& a, unsigned sb, unsigned char* lds)
{
	const uint64_t r0 = a.clock->stdp_first;
	const unsigned n = unsigned(a.clock->stdp_count);
	const unsigned cur_slot = unsigned(a.clock->stdp_snaps & 1);
	if(n == 0 || n > 64) return;
	const nemo_hip_stdp& st = a.stdp;
	int32_t* s_fn = reinterpret_cast<int32_t*>(lds);
	uint32_t* s_rb = reinterpret_cast<uint32_t*>(lds + 128 * sizeof(int32_t));
	uint32_t* s_pos = s_rb + THREADS;
	uint64_t* s_post = reinterpret_cast<uint64_t*>(s_pos + THREADS + 2);
	const unsigned P = st.post_window;
	const unsigned tid = threadIdx.x;
	if(tid < 64) {
		s_fn[tid] = st.fn_pre[tid];
		s_fn[64 + tid] = st.fn_post[tid];
	}
	const uint64_t rl = r0 + n - 1;
	/* the two snapshots, after rl and after r0 - 1: one 16-byte load per
	 * neuron */
	const ulonglong2* rec = reinterpret_cast<const ulonglong2*>(a.stdp_snap);
	auto split = [cur_slot](const ulonglong2 v, uint64_t& cur, uint64_t& old) {
		cur = cur_slot ? v.y : v.x;
		old = cur_slot ? v.x : v.y;
	};
	/* bit y: step rl - y is accumulated (in the batch, r >= P, not flushed) */
	uint64_t due = n >= 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1;
	if(rl < P) {
		due = 0;
	} else if(rl - P < 63) {
		due &= (uint64_t(2) << (rl - P)) - 1;
	}
	const uint64_t flushed = a.clock->stdp_flushed;
	if(flushed >= r0 && flushed <= rl) due &= ~(uint64_t(1) << (rl - flushed));

	/* the block's words, THREADS / 64 (one target per thread) at a time */
	unsigned long long examined_all = 0;
	for(unsigned wg = 0; wg < a.stdp_words; wg += THREADS / 64) {
		__syncthreads(); /* the previous group is done with the LDS lists */
		const size_t w0 = size_t(sb) * a.stdp_words + wg;
		const size_t word = w0 + tid / 64;
		uint64_t post = 0; /* bit y: the target fired at step rl - y - P */
		if(tid / 64 < min(THREADS / 64, a.stdp_words - wg) && word < a.fired_words) {
			const uint32_t tn = uint32_t(w0 * 64 + tid);
			uint64_t c, o;
			split(rec[tn], c, o);
			post = history_window(c, o, n, P) & due;
		}
		const bool fired = post != 0;
		uint32_t nfired;
		const uint32_t idx = block_exclusive_scan<THREADS>(fired ? 1u : 0u, &nfired);
		uint32_t len = 0, rb = 0;
		if(fired) {
			const uint32_t tn = uint32_t(w0 * 64 + tid);
			rb = st.rcm_ptr[tn];
			len = st.rcm_ptr[tn + 1] - rb;
		}
		uint32_t total;
		const uint32_t pos = block_exclusive_scan<THREADS>(len, &total);
		uint32_t examined;
		block_exclusive_scan<THREADS>(len * unsigned(__popcll(post)), &examined);
		examined_all += examined;
		if(fired) {
			s_rb[idx] = rb;
			s_pos[idx] = pos;
			s_post[idx] = post;
		}
		if(tid == 0) s_pos[nfired] = total;
		__syncthreads();
		if(total == 0) continue;

		for(uint32_t e0 = tid; e0 < total; e0 += SYN_ILP * THREADS) {
			uint32_t k[SYN_ILP];
			uint32_t sd[SYN_ILP];
			int32_t acc[SYN_ILP];
			uint64_t pm[SYN_ILP];
	#pragma unroll
			for(int u = 0; u < SYN_ILP; ++u) k[u] = e0 + u * THREADS;
			unsigned lo[SYN_ILP];
			last_not_above_k<THREADS>(s_pos, nfired, k, lo);
	#pragma unroll
			for(int u = 0; u < SYN_ILP; ++u) {
				const uint32_t e = k[u];
				k[u] = 0;
				sd[u] = 0;
				acc[u] = 0;
				pm[u] = 0;
				if(e < total) {
					k[u] = s_rb[lo[u]] + (e - s_pos[lo[u]]);
					pm[u] = s_post[lo[u]];
					sd[u] = st.rcm_sd[k[u]];
					acc[u] = st.rcm_acc[k[u]];
				}
			}
			uint64_t hc[SYN_ILP], hs[SYN_ILP];
	#pragma unroll
			for(int u = 0; u < SYN_ILP; ++u) {
				hc[u] = 0;
				hs[u] = 0;
				if(e0 + u * THREADS < total) {
					split(rec[sd[u] & 0x3ffffffu], hc[u], hs[u]);
				}
			}
	#pragma unroll
			for(int u = 0; u < SYN_ILP; ++u) {
				if(e0 + u * THREADS < total) {
					const unsigned d = (sd[u] >> 26) + 1;
					int32_t dw = 0;
					for(uint64_t m = pm[u]; m; m &= m - 1) {
						const unsigned y = unsigned(__builtin_ctzll(m));
						const uint64_t h = history_window(hc[u], hs[u], n, y);
						const uint64_t preFiring = d >= 64 ? 0 : (h >> d);
						dw = stdp_wrap_add(dw, stdp_wrap_add(
								stdp_update_region(preFiring & st.potentiation, P, st.pre_post, st.post_pre,
									s_fn, s_fn + 64),
								stdp_update_region(preFiring & st.depression, P, st.pre_post, st.post_pre,
									s_fn, s_fn + 64)));
					}
					if(dw != 0) {
						st.rcm_acc[k[u]] = stdp_wrap_add(acc[u], dw);
					}
				}
			}
		}
	}
	if(tid == 0) atomicAdd(&st.examined[sb], examined_all);
}

/* STDP accumulation of step t-1 by STDP block sb, unless it is not due */
template<unsigned THREADS>
__device__ inline void stdp_step(const nemo_propagate_args& a, uint64_t t, unsigned sb, unsigned char* lds)
{
	if(t == 0) return;
	const uint64_t r = t - 1;
	if(r == a.clock->stdp_flushed || r < a.stdp.post_window) return;
	stdp_block<THREADS>(a, r, sb, lds);
}


/* In an update launch of NEMO_COLAUNCH_BLOCK threads per block: run this
 * block as one of the propagate blocks the launch carries after the group's
 * tiles (nemo_hip_colaunch) and return true, or, in a block of the group's
 * own tiles, return false for its first NEMO_TILE threads (the tile's) and,
 * for the others, true (their waves then end; a block barrier waits only for
 * the waves still running) or, with GATHER_HELPERS, false: the kernel then
 * passes them to gather_input<NEMO_COLAUNCH_BLOCK>, which releases them.
 * Every update kernel of such a launch calls it first and returns when it
 * returns true. */
template<bool GATHER_HELPERS = false>
__device__ inline bool colaunched_block(const nemo_hip_step_buffers& b, const nemo_hip_neuron_group& g)
{
	static_assert(NEMO_COLAUNCH_BLOCK % 64 == 0 && NEMO_COLAUNCH_BLOCK >= NEMO_TILE, "whole waves beyond the tile");
	const unsigned tiles = (g.count + NEMO_TILE - 1) / NEMO_TILE;
	if(blockIdx.x < tiles) return !GATHER_HELPERS && threadIdx.x >= NEMO_TILE;
	extern __shared__ uint64_t colaunch_lds[];
	unsigned char* lds = reinterpret_cast<unsigned char*>(colaunch_lds);
	const unsigned k = blockIdx.x - tiles;
	const nemo_propagate_args& a = b.co.args;
	const uint64_t t = a.clock->base_step + a.step_offset;
	if(k < b.co.deliver_blocks) {
		deliver_block<NEMO_COLAUNCH_BLOCK, CARRIED_EVENT_ILP>(a, t, k, lds);
	} else {
		stdp_step<NEMO_COLAUNCH_BLOCK>(a, t - 1, k - b.co.deliver_blocks, lds);
	}
	return true;
}

} // namespace nemo_dev

#endif
