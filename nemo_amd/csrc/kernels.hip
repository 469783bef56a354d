/*
 * The backend's own gfx950 kernels: the per-step propagate kernel (binned
 * spike delivery), the STDP-only kernel, applyStdp and stimulus helpers.
 * Neuron updates, which also gather the binned input and usually run the
 * STDP accumulation, live in the model plugins (plugins/<Model>.hip).
 *
 * Data layout in HBM.  Device indices number neurons type by type, in
 * insertion order within a type (NeMo's simulation order,
 * src/nemo/cpu/Simulation.cpp:60-77), each type starting at a multiple of
 * NEMO_TILE (256):
 *   fcm        int2[S]     synapse rows (source, delay) back to back; row
 *                          (n, d) = [rowptr[n*D+d-1], rowptr[n*D+d]); entries
 *                          {target, Q11.20 weight}, 8 bytes, so consecutive
 *                          lanes read consecutive 8-byte words of a row
 *   fired_bits uint64[128][N/64]  one firing bitmap per step (ring)
 *   hist       uint64[2][N] firing history (bit k: fired k steps ago); plane
 *                          s % 2 holds the history after step s
 *   bins       int2[2][B][CAP] each delivery block's events, sorted by tile,
 *                          for steps of even and odd parity
 *   bin_table  uint32[2][T][B] count << 16 | offset of tile t in block b's
 *                          bins (entries of the blocks whose window holds t)
 *   acc        int64[2][2][N] overflow accumulators (events beyond CAP)
 *   rcm_*      per target, its plastic incoming synapses (STDP only)
 */
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "glibc_math.h"
#include "plugin_device.h"

namespace {

constexpr uint64_t RING_MASK = NEMO_STEP_RING - 1;
/* fired rows a delivery block lists in LDS at a time */
constexpr unsigned ROWS = 1024;
constexpr unsigned ROWS_PER_THREAD = ROWS / NEMO_BLOCK;
/* independent global loads each thread keeps in flight in the synapse loop
 * of STDP, and in the event loop of delivery */
constexpr int ILP = 8;
constexpr int EVENT_ILP = 16;

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
__device__ void deliver_block(const nemo_propagate_args& a, uint64_t t, int2* stage, uint32_t* cnt, uint32_t* cur)
{
	__shared__ uint64_t s_word[NEMO_BLOCK];
	__shared__ uint32_t s_wpre[NEMO_BLOCK + 1];
	__shared__ uint32_t s_rb[ROWS];
	__shared__ uint32_t s_rpre[ROWS + 1];

	const unsigned CAP = a.bin_capacity;
	const unsigned D = a.D;
	const unsigned DC = a.d_count;
	const unsigned tid = threadIdx.x;
	const unsigned b = blockIdx.x;
	const unsigned col = a.block_base + b;
	const size_t parity = t & 1;
	int64_t* acc = a.acc + parity * 2 * a.acc_pitch;
	const uint2 win = a.windows[b];
	const unsigned tlo = win.x;
	const unsigned W = win.y - win.x; /* tiles in the window */
	for(unsigned p = tid; p < W; p += NEMO_BLOCK) cnt[p] = 0;

	const uint64_t q0 = a.pair_bounds[b];
	const uint64_t q1 = a.pair_bounds[b + 1];
	const uint64_t nwords = (q1 - q0) * 4;
	uint32_t fill = 0;
	unsigned long long events = 0;

	for(uint64_t wb = 0; wb < nwords; wb += NEMO_BLOCK) {
		/* the bitmap words of this batch, one per thread */
		uint64_t word = 0;
		const uint64_t w = wb + tid;
		if(w < nwords) {
			const uint64_t q = q0 + w / 4;
			const unsigned p = unsigned(q / DC);
			const unsigned d = unsigned(q % DC) + a.d_first;
			if(t >= d) word = a.fired_bits[((t - d) & RING_MASK) * a.fired_words + size_t(p) * 4 + (w & 3)];
		}
		__syncthreads(); /* previous batch done with s_word */
		s_word[tid] = word;
		uint32_t R;
		const uint32_t wpre = nemo_dev::block_exclusive_scan(uint32_t(__popcll(word)), &R);
		s_wpre[tid] = wpre;
		if(tid == 0) s_wpre[NEMO_BLOCK] = R;
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
			nemo_dev::last_not_above_k<NEMO_BLOCK>(s_wpre, NEMO_BLOCK, rr, wl);
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
			uint32_t pre = nemo_dev::block_exclusive_scan(mine, &M);
#pragma unroll
			for(unsigned k = 0; k < ROWS_PER_THREAD; ++k) {
				const uint32_t ri = tid * ROWS_PER_THREAD + k;
				if(ri < nr) s_rpre[ri] = pre;
				pre += lens[k];
			}
			if(tid == 0) s_rpre[nr] = M;
			__syncthreads();

			/* the events of these rows, evenly over the block */
			for(uint32_t e0 = tid; e0 < M; e0 += EVENT_ILP * NEMO_BLOCK) {
				uint32_t ee[EVENT_ILP];
				unsigned lo[EVENT_ILP];
#pragma unroll
				for(int u = 0; u < EVENT_ILP; ++u) ee[u] = e0 + u * NEMO_BLOCK;
				nemo_dev::last_not_above_k<ROWS>(s_rpre, nr, ee, lo);
				int2 s[EVENT_ILP];
#pragma unroll
				for(int u = 0; u < EVENT_ILP; ++u) {
					s[u] = make_int2(0, 0);
					if(ee[u] < M) s[u] = a.fcm[s_rb[lo[u]] + (ee[u] - s_rpre[lo[u]])];
				}
#pragma unroll
				for(int u = 0; u < EVENT_ILP; ++u) {
					const uint32_t e = e0 + u * NEMO_BLOCK;
					if(e < M) {
						const uint32_t pos = fill + e;
						const uint32_t rel = uint32_t(s[u].x) / NEMO_TILE - tlo;
						if(pos < CAP && rel < W) {
							stage[pos] = s[u];
							atomicAdd(&cnt[rel], 1u);
						} else {
							/* beyond capacity (or, defensively, outside the
							 * window): straight into the accumulators */
							if(pos < CAP) stage[pos] = make_int2(-1, 0);
							int64_t* dst = acc + (s[u].y >= 0 ? size_t(0) : a.acc_pitch) + uint32_t(s[u].x);
							atomicAdd(reinterpret_cast<unsigned long long*>(dst),
									(unsigned long long)(int64_t)s[u].y);
						}
					}
				}
			}
			fill += M;
			events += M;
			__syncthreads(); /* s_rb / s_rpre are rewritten by the next chunk */
		}
	}

	/* counting sort of the staged events by target tile */
	const uint32_t staged = min(fill, CAP);
	uint32_t carry = 0;
	for(uint32_t base = 0; base < W; base += NEMO_BLOCK) {
		const uint32_t p = base + tid;
		const uint32_t c = p < W ? cnt[p] : 0u;
		uint32_t total;
		const uint32_t pre = nemo_dev::block_exclusive_scan(c, &total);
		if(p < W) {
			cur[p] = carry + pre;
			a.bin_table[(parity * a.tiles + tlo + p) * a.table_columns + col] = (c << 16) | (carry + pre);
		}
		carry += total;
	}
	__syncthreads();
	int2* out = a.bins + (parity * a.table_columns + col) * CAP;
	/* four staged events per thread at a time, so that their LDS atomics
	 * (which return the slot) overlap */
	for(uint32_t i0 = tid; i0 < staged; i0 += 4 * NEMO_BLOCK) {
		int2 e[4];
		uint32_t pos[4];
#pragma unroll
		for(int u = 0; u < 4; ++u) {
			const uint32_t i = i0 + u * NEMO_BLOCK;
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
	if(tid == 0 && events) a.block_events[col] += events;
}

/* ---------------------------------------------------------------- STDP */

using nemo_dev::stdp_wrap_add;

/* An STDP block of the STDP-only launch: the targets among its bitmap words
 * that fired P steps before step r, and all their plastic incoming synapses
 * spread evenly over its threads (ILP loads in flight per thread).  The
 * update kernels do the same per tile when STDP runs inside them. */
__device__ void stdp_block(const nemo_propagate_args& a, uint64_t r, unsigned sb)
{
	const nemo_hip_stdp& st = a.stdp;
	__shared__ int32_t s_fn[128];
	__shared__ uint32_t s_rb[NEMO_BLOCK];
	__shared__ uint32_t s_pos[NEMO_BLOCK + 1];
	const unsigned P = st.post_window;
	const unsigned tid = threadIdx.x;
	if(tid < 64) {
		s_fn[tid] = st.fn_pre[tid];
		s_fn[64 + tid] = st.fn_post[tid];
	}

	const uint64_t slot = (r - P) & RING_MASK;
	const size_t w0 = size_t(sb) * a.stdp_words;
	const size_t word = w0 + tid / 64;
	bool fired = false;
	if(tid / 64 < a.stdp_words && word < a.fired_words) {
		fired = (a.fired_bits[slot * a.fired_words + word] >> (tid & 63)) & 1u;
	}
	uint32_t nfired;
	const uint32_t idx = nemo_dev::block_exclusive_scan(fired ? 1u : 0u, &nfired);
	uint32_t len = 0, rb = 0;
	if(fired) {
		const uint32_t n = uint32_t(w0 * 64 + tid);
		rb = st.rcm_ptr[n];
		len = st.rcm_ptr[n + 1] - rb;
	}
	uint32_t total;
	const uint32_t pos = nemo_dev::block_exclusive_scan(len, &total);
	if(fired) {
		s_rb[idx] = rb;
		s_pos[idx] = pos;
	}
	if(tid == 0) s_pos[nfired] = total;
	__syncthreads();
	if(total == 0) return;

	const uint64_t* hist = a.hist + (r & 1) * a.hist_pitch;
	for(uint32_t e0 = tid; e0 < total; e0 += ILP * NEMO_BLOCK) {
		uint32_t k[ILP];
		uint32_t sd[ILP];
		int32_t acc[ILP];
#pragma unroll
		for(int u = 0; u < ILP; ++u) k[u] = e0 + u * NEMO_BLOCK;
		unsigned lo[ILP];
		nemo_dev::last_not_above_k<NEMO_BLOCK>(s_pos, nfired, k, lo);
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
			if(e0 + u * NEMO_BLOCK < total) h[u] = hist[sd[u] & 0x3ffffffu];
		}
#pragma unroll
		for(int u = 0; u < ILP; ++u) {
			if(e0 + u * NEMO_BLOCK < total) {
				const unsigned d = (sd[u] >> 26) + 1;
				const uint64_t preFiring = d >= 64 ? 0 : (h[u] >> d);
				const int32_t dw = stdp_wrap_add(
						nemo_dev::stdp_update_region(preFiring & st.potentiation, P, st.pre_post, st.post_pre,
							s_fn, s_fn + 64),
						nemo_dev::stdp_update_region(preFiring & st.depression, P, st.pre_post, st.post_pre,
							s_fn, s_fn + 64));
				if(dw != 0) {
					st.rcm_acc[k[u]] = stdp_wrap_add(acc[u], dw);
				}
			}
		}
	}
	if(tid == 0) st.examined[sb] += total;
}

/* STDP accumulation of step t-1 by STDP block sb, unless it is not due */
__device__ inline void stdp_step(const nemo_propagate_args& a, uint64_t t, unsigned sb)
{
	if(t == 0) return;
	const uint64_t r = t - 1;
	if(r == a.clock->stdp_flushed || r < a.stdp.post_window) return;
	stdp_block(a, r, sb);
}

/* delivery blocks, then (when fused into this launch) STDP blocks */
__global__ void __launch_bounds__(NEMO_BLOCK)
propagate_kernel(nemo_propagate_args a)
{
	extern __shared__ int2 lds_dyn[];
	const uint64_t t = a.clock->base_step + a.step_offset;
	if(blockIdx.x < a.deliver_blocks) {
		int2* stage = lds_dyn;
		uint32_t* cnt = reinterpret_cast<uint32_t*>(lds_dyn + a.bin_capacity);
		deliver_block(a, t, stage, cnt, cnt + a.max_window);
	} else {
		stdp_step(a, t, blockIdx.x - a.deliver_blocks);
	}
}

/* STDP blocks only */
__global__ void __launch_bounds__(NEMO_BLOCK)
stdp_kernel(nemo_propagate_args a)
{
	stdp_step(a, a.clock->base_step + a.step_offset, blockIdx.x);
}

__global__ void __launch_bounds__(NEMO_BLOCK)
apply_stdp_kernel(nemo_apply_args a)
{
	for(size_t k = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x; k < a.count;
			k += size_t(gridDim.x) * NEMO_BLOCK) {
		const int32_t acc = a.reward == 0 ? 0 : a.rcm_acc[k];
		int32_t* wp = reinterpret_cast<int32_t*>(a.fcm + a.rcm_fwd[k]) + 1;
		const int32_t w_old = *wp;
		const int32_t dw = int32_t((int64_t(a.reward) * int64_t(acc)) >> a.fbits);
		int32_t w_new = 0;
		if(w_old > 0) {
			const int32_t x = stdp_wrap_add(w_old, dw);
			w_new = min(a.maxE, max(x, a.minE));
		} else if(w_old < 0) {
			const int32_t x = stdp_wrap_add(w_old, dw);
			w_new = min(a.minI, max(x, a.maxI));
		}
		if(w_new != w_old) *wp = w_new;
		if(a.rcm_acc[k] != 0) a.rcm_acc[k] = 0;
	}
}

/* ------------------------------------------------------------ stimulus */

__global__ void fstim_kernel(uint32_t* bits, const uint32_t* idx, size_t n, int set)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) {
		const uint32_t v = idx[i];
		if(set) atomicOr(&bits[v >> 5], 1u << (v & 31));
		else atomicAnd(&bits[v >> 5], ~(1u << (v & 31)));
	}
}

__global__ void istim_kernel(float* ext, const uint32_t* idx, const float* val, size_t n)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) ext[idx[i]] = val[i];
}

/* plain streaming copy, 16 bytes per thread per iteration */
__global__ void __launch_bounds__(NEMO_BLOCK)
copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, size_t n)
{
	for(size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x; i < n; i += size_t(gridDim.x) * NEMO_BLOCK) {
		dst[i] = src[i];
	}
}

__global__ void eval_math_kernel(int fn, const float* in, float* out, size_t n)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) {
		const float x = in[i];
		out[i] = fn == 0 ? nemo_math::logf_exact(x) : fn == 1 ? nemo_math::sinf_exact(x) : nemo_math::expf_exact(x);
	}
}

} // namespace

extern "C" {

size_t nemo_propagate_lds_bytes(unsigned max_window, unsigned bin_capacity)
{
	return size_t(bin_capacity) * sizeof(int2) + 2 * size_t(max_window) * sizeof(uint32_t);
}

hipError_t nemo_launch_propagate(hipStream_t st, const nemo_propagate_args* a, unsigned stdp_blocks, int fuse_stdp)
{
	if(a->deliver_blocks > NEMO_MAX_BIN_BLOCKS) return hipErrorInvalidValue;
	if(a->deliver_blocks && (a->D == 0 || a->D > NEMO_MAX_DELAY || a->bin_capacity == 0
				|| a->bin_capacity > 32768 || a->max_window == 0 || a->max_window > a->tiles
				|| !a->pair_bounds || !a->windows || a->d_first == 0 || a->d_count == 0
				|| a->d_first + a->d_count - 1 > a->D || a->block_base + a->deliver_blocks > a->table_columns
				|| a->table_columns > NEMO_MAX_BIN_BLOCKS)) {
		return hipErrorInvalidValue;
	}
	if(stdp_blocks && (a->stdp.post_window > 63 || a->stdp_words == 0 || a->stdp_words > 4
				|| size_t(stdp_blocks) * a->stdp_words < a->fired_words || !a->stdp.rcm_ptr || !a->hist)) {
		return hipErrorInvalidValue;
	}
	const size_t lds = a->deliver_blocks ? nemo_propagate_lds_bytes(a->max_window, a->bin_capacity) : 0;
	if(fuse_stdp) {
		const unsigned blocks = a->deliver_blocks + stdp_blocks;
		if(blocks == 0) return hipSuccess;
		hipLaunchKernelGGL(propagate_kernel, dim3(blocks), dim3(NEMO_BLOCK), lds, st, *a);
		return hipGetLastError();
	}
	if(stdp_blocks) {
		hipLaunchKernelGGL(stdp_kernel, dim3(stdp_blocks), dim3(NEMO_BLOCK), 0, st, *a);
		const hipError_t e = hipGetLastError();
		if(e != hipSuccess) return e;
	}
	if(a->deliver_blocks) {
		hipLaunchKernelGGL(propagate_kernel, dim3(a->deliver_blocks), dim3(NEMO_BLOCK), lds, st, *a);
		return hipGetLastError();
	}
	return hipSuccess;
}

hipError_t nemo_launch_apply_stdp(hipStream_t st, const nemo_apply_args* a)
{
	if(a->count == 0) return hipSuccess;
	size_t blocks = (a->count + NEMO_BLOCK - 1) / NEMO_BLOCK;
	if(blocks > 8192) blocks = 8192;
	hipLaunchKernelGGL(apply_stdp_kernel, dim3(unsigned(blocks)), dim3(NEMO_BLOCK), 0, st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_fstim(hipStream_t st, uint32_t* bits, const uint32_t* idx, size_t n, int set)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(fstim_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			bits, idx, n, set);
	return hipGetLastError();
}

hipError_t nemo_launch_istim(hipStream_t st, float* ext, const uint32_t* idx, const float* val, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(istim_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			ext, idx, val, n);
	return hipGetLastError();
}

hipError_t nemo_launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes)
{
	const size_t n = bytes / sizeof(int4);
	if(n == 0) return hipSuccess;
	size_t blocks = (n + NEMO_BLOCK - 1) / NEMO_BLOCK;
	if(blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(copy_kernel, dim3(unsigned(blocks)), dim3(NEMO_BLOCK), 0, st,
			static_cast<const int4*>(src), static_cast<int4*>(dst), n);
	return hipGetLastError();
}

hipError_t nemo_launch_eval_math(hipStream_t st, int fn, const float* in, float* out, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(eval_math_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			fn, in, out, n);
	return hipGetLastError();
}

} // extern "C"
