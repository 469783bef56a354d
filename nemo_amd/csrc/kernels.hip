/*
 * The backend's own gfx950 kernels: spike delivery, STDP accumulation,
 * applyStdp, the step clock and stimulus helpers.  Neuron updates live in the
 * model plugins (plugins/*.hip).
 *
 * Data layout in HBM (all indices are simulation indices: neurons numbered
 * type by type, in insertion order within a type, as NeMo's CPU backend does,
 * src/nemo/cpu/Simulation.cpp:60-77):
 *   fcm      int2[S]       synapse rows (source, delay) back to back; row
 *                          (n, d) = [rowptr[n*D+d-1], rowptr[n*D+d]); each
 *                          entry {target, Q11.20 weight}: 8 bytes, so a wave
 *                          reads 512 contiguous bytes per row chunk
 *   acc      int64[2][N]   excitatory / inhibitory current accumulators
 *   hist     uint64[N]     firing history (bit k: fired k steps ago)
 *   fired_log uint32[2^k]  circular log of fired neurons; step s's neurons are
 *                          log[step_start[s] .. step_start[s+1])
 *   rcm_*    per target, its plastic incoming synapses (STDP only)
 */
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "glibc_math.h"

#define NEMO_BLOCK 256

namespace {

constexpr uint64_t RING_MASK = NEMO_STEP_RING - 1;

/* ------------------------------------------------------------ delivery */

template<int G>
__global__ void __launch_bounds__(NEMO_BLOCK)
deliver_kernel(nemo_deliver_args a)
{
	__shared__ uint64_t s_begin[NEMO_MAX_DELAY];
	__shared__ uint32_t s_prefix[NEMO_MAX_DELAY + 1];
	__shared__ unsigned long long s_events;

	const uint64_t t = a.clock->step;
	const unsigned D = a.D;

	/* wave 0: the fired lists of steps t-1 .. t-D and an exclusive scan of
	 * their lengths, so that work item j maps to (delay bin, list position) */
	if(threadIdx.x < 64) {
		const unsigned d = threadIdx.x + 1;
		uint64_t beg = 0;
		uint32_t cnt = 0;
		if(d <= D && t >= d) {
			const uint64_t s = t - d;
			beg = a.step_start[s & RING_MASK];
			cnt = uint32_t(a.step_start[(s + 1) & RING_MASK] - beg);
		}
		s_begin[threadIdx.x] = beg;
		uint32_t x = cnt;
#pragma unroll
		for(int o = 1; o < 64; o <<= 1) {
			const uint32_t y = __shfl_up(x, o);
			if(int(threadIdx.x) >= o) x += y;
		}
		s_prefix[threadIdx.x + 1] = x;
		if(threadIdx.x == 0) {
			s_prefix[0] = 0;
			s_events = 0;
		}
	}
	__syncthreads();

	const uint32_t total = s_prefix[D];
	constexpr unsigned GROUPS = NEMO_BLOCK / G;
	const unsigned lane = threadIdx.x % G;
	unsigned long long events = 0;

	for(uint32_t j = blockIdx.x * GROUPS + threadIdx.x / G; j < total; j += gridDim.x * GROUPS) {
		unsigned lo = 0, hi = D; /* s_prefix[lo] <= j < s_prefix[hi] */
		while(hi - lo > 1) {
			const unsigned mid = (lo + hi) >> 1;
			if(s_prefix[mid] <= j) lo = mid; else hi = mid;
		}
		const uint32_t n = a.fired_log[(s_begin[lo] + (j - s_prefix[lo])) & a.log_mask];
		const size_t row = size_t(n) * D + lo;
		const uint32_t b = a.rowptr[row];
		const uint32_t e = a.rowptr[row + 1];
		if(lane == 0) events += e - b;
		/* four independent loads in flight per lane before the atomics */
		for(uint32_t k0 = b + lane; k0 < e; k0 += 4 * G) {
			int2 s[4];
#pragma unroll
			for(int u = 0; u < 4; ++u) {
				const uint32_t k = k0 + u * G;
				s[u] = k < e ? a.fcm[k] : make_int2(0, 0);
			}
#pragma unroll
			for(int u = 0; u < 4; ++u) {
				if(k0 + u * G < e) {
					int64_t* p = a.acc + (s[u].y >= 0 ? size_t(0) : a.acc_pitch) + uint32_t(s[u].x);
					atomicAdd(reinterpret_cast<unsigned long long*>(p),
							(unsigned long long)(int64_t)s[u].y);
				}
			}
		}
	}
	if(events) atomicAdd(&s_events, events);
	__syncthreads();
	if(threadIdx.x == 0 && s_events) {
		atomicAdd(reinterpret_cast<unsigned long long*>(&a.clock->events), s_events);
	}
}

/* ---------------------------------------------------------------- STDP */

__device__ inline int32_t wrap_add(int32_t x, int32_t y)
{
	return int32_t(uint32_t(x) + uint32_t(y));
}

__device__ inline int32_t update_region(uint64_t arrivals, const nemo_stdp_args& a,
		const int32_t* fnPre, const int32_t* fnPost)
{
	/* StdpProcess::updateRegion, StdpProcess.cpp:86-111: the nearest arrival
	 * on either side of the postsynaptic spike wins; a tie gives nothing. */
	if(!arrivals) return 0;
	const unsigned NONE = ~0u;
	const uint64_t pre = arrivals & a.prePostBits;
	const uint64_t post = arrivals & a.postPreBits;
	const unsigned dt_pre = pre ? unsigned(__builtin_ctzll(pre >> a.postPreWindow)) : NONE;
	const unsigned dt_post = post ? unsigned(__builtin_clzll(post << (64 - a.postPreWindow))) : NONE;
	if(dt_pre < dt_post) return fnPre[dt_pre];
	if(dt_post < dt_pre) return fnPost[dt_post];
	return 0;
}

template<int G>
__global__ void __launch_bounds__(NEMO_BLOCK)
stdp_kernel(nemo_stdp_args a)
{
	__shared__ int32_t s_pre[64];
	__shared__ int32_t s_post[64];
	__shared__ unsigned long long s_examined;
	const uint64_t t = a.clock->step;
	const unsigned P = a.postPreWindow;
	if(t < P) return;
	if(threadIdx.x < 64) {
		s_pre[threadIdx.x] = a.fnPre[threadIdx.x];
		s_post[threadIdx.x] = a.fnPost[threadIdx.x];
	}
	if(threadIdx.x == 0) s_examined = 0;
	__syncthreads();
	unsigned long long examined = 0;

	/* the targets whose firing history has bit P set: those that fired at t-P */
	const uint64_t s = t - P;
	const uint64_t beg = a.step_start[s & RING_MASK];
	const uint64_t end = P == 0 ? a.clock->log_head : a.step_start[(s + 1) & RING_MASK];
	const uint32_t total = uint32_t(end - beg);
	constexpr unsigned GROUPS = NEMO_BLOCK / G;
	const unsigned lane = threadIdx.x % G;

	for(uint32_t j = blockIdx.x * GROUPS + threadIdx.x / G; j < total; j += gridDim.x * GROUPS) {
		const uint32_t n = a.fired_log[(beg + j) & a.log_mask];
		const uint32_t rb = a.rcm_ptr[n];
		const uint32_t re = a.rcm_ptr[n + 1];
		if(lane == 0) examined += re - rb;
		for(uint32_t k = rb + lane; k < re; k += G) {
			const uint32_t sd = a.rcm_sd[k];
			const uint64_t preFiring = a.hist[sd & 0x3ffffffu] >> ((sd >> 26) + 1);
			const int32_t dw = wrap_add(update_region(preFiring & a.potentiationBits, a, s_pre, s_post),
					update_region(preFiring & a.depressionBits, a, s_pre, s_post));
			if(dw != 0) {
				a.rcm_acc[k] = wrap_add(a.rcm_acc[k], dw);
			}
		}
	}
	if(examined) atomicAdd(&s_examined, examined);
	__syncthreads();
	if(threadIdx.x == 0 && s_examined) {
		atomicAdd(reinterpret_cast<unsigned long long*>(&a.clock->stdp_synapses), s_examined);
	}
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
			const int32_t x = wrap_add(w_old, dw);
			w_new = min(a.maxE, max(x, a.minE));
		} else if(w_old < 0) {
			const int32_t x = wrap_add(w_old, dw);
			w_new = min(a.minI, max(x, a.maxI));
		}
		if(w_new != w_old) *wp = w_new;
		if(a.rcm_acc[k] != 0) a.rcm_acc[k] = 0;
	}
}

/* ------------------------------------------------------------ bookkeeping */

__global__ void advance_kernel(nemo_hip_clock* c, uint64_t* step_start)
{
	if(threadIdx.x == 0 && blockIdx.x == 0) {
		const uint64_t next = c->step + 1;
		c->step = next;
		c->cycle = c->cycle + 1;
		step_start[next & RING_MASK] = c->log_head;
	}
}

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

__global__ void eval_math_kernel(int fn, const float* in, float* out, size_t n)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) {
		const float x = in[i];
		out[i] = fn == 0 ? nemo_math::logf_exact(x) : fn == 1 ? nemo_math::sinf_exact(x) : nemo_math::expf_exact(x);
	}
}

template<int G>
hipError_t launchDeliverG(hipStream_t st, const nemo_deliver_args* a, unsigned blocks)
{
	hipLaunchKernelGGL(deliver_kernel<G>, dim3(blocks), dim3(NEMO_BLOCK), 0, st, *a);
	return hipGetLastError();
}

template<int G>
hipError_t launchStdpG(hipStream_t st, const nemo_stdp_args* a, unsigned blocks)
{
	hipLaunchKernelGGL(stdp_kernel<G>, dim3(blocks), dim3(NEMO_BLOCK), 0, st, *a);
	return hipGetLastError();
}

} // namespace

extern "C" {

hipError_t nemo_launch_deliver(hipStream_t st, const nemo_deliver_args* a, unsigned G, unsigned blocks)
{
	if(a->D == 0 || a->D > NEMO_MAX_DELAY || blocks == 0) return hipErrorInvalidValue;
	switch(G) {
		case 1: return launchDeliverG<1>(st, a, blocks);
		case 2: return launchDeliverG<2>(st, a, blocks);
		case 4: return launchDeliverG<4>(st, a, blocks);
		case 8: return launchDeliverG<8>(st, a, blocks);
		case 16: return launchDeliverG<16>(st, a, blocks);
		case 32: return launchDeliverG<32>(st, a, blocks);
		case 64: return launchDeliverG<64>(st, a, blocks);
		default: return hipErrorInvalidValue;
	}
}

hipError_t nemo_launch_stdp(hipStream_t st, const nemo_stdp_args* a, unsigned G, unsigned blocks)
{
	if(a->postPreWindow > 63 || blocks == 0) return hipErrorInvalidValue;
	switch(G) {
		case 1: return launchStdpG<1>(st, a, blocks);
		case 2: return launchStdpG<2>(st, a, blocks);
		case 4: return launchStdpG<4>(st, a, blocks);
		case 8: return launchStdpG<8>(st, a, blocks);
		case 16: return launchStdpG<16>(st, a, blocks);
		case 32: return launchStdpG<32>(st, a, blocks);
		case 64: return launchStdpG<64>(st, a, blocks);
		default: return hipErrorInvalidValue;
	}
}

hipError_t nemo_launch_apply_stdp(hipStream_t st, const nemo_apply_args* a)
{
	if(a->count == 0) return hipSuccess;
	size_t blocks = (a->count + NEMO_BLOCK - 1) / NEMO_BLOCK;
	if(blocks > 8192) blocks = 8192;
	hipLaunchKernelGGL(apply_stdp_kernel, dim3(unsigned(blocks)), dim3(NEMO_BLOCK), 0, st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_advance(hipStream_t st, nemo_hip_clock* c, uint64_t* step_start)
{
	hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, st, c, step_start);
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

hipError_t nemo_launch_eval_math(hipStream_t st, int fn, const float* in, float* out, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(eval_math_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			fn, in, out, n);
	return hipGetLastError();
}

} // extern "C"
