/*
 * Kuramoto oscillator plugin (libKuramoto.so), bit-exact with
 * src/nemo/cpu/plugins/Kuramoto.cpp:14-159; built with NEMO_KURAMOTO_LAG
 * defined (plugins/KuramotoLag.hip) it is libKuramotoLag.so, bit-exact with
 * src/nemo/cpu/plugins/KuramotoLag.cpp:15-138: a third parameter alpha is
 * subtracted inside every sine, sin(theta_k - theta - alpha).
 *
 * Parameters (plugins/Kuramoto.h): omega, the phase advance per unit time,
 * and Cmean, the coupling normalisation (1 when not positive).  State: the
 * phase theta, with a history of H = 32 slots (Kuramoto.ini): the update
 * at cycle c reads slot c % H and writes slot (c + 1) % H.  Per step, one
 * RK4 step of h = 0.05 of
 *   dtheta/dt = omega + sum_k w_k sin(theta_k(c - (d_k - 1)) - theta) / Cmean
 * over the oscillator's incoming synapses k (source phase d_k - 1 cycles
 * back), summed in the order of the reverse matrix; the new phase is
 * fmodf(theta, float(2 pi)), plus 2 pi in double when theta < 0.
 * Oscillators never fire and read no synaptic current.  A synapse whose
 * delay reads the history slot the step writes (delay 32 or 64) sees what
 * the CPU backend's neuron-by-neuron loop sees: the source's new phase when
 * the source's index is lower, its old phase otherwise; the backend then
 * orders the updates in passes (nemo_hip_neuron_group.levels).
 *
 * The sine is glibc's sinf restated (glibc_math.h), which this library
 * evaluates for |x| < 120; a larger argument sets NEMO_HIP_STATUS_SINF_RANGE
 * and the host reports it.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

enum { PARAM_FREQUENCY = 0, PARAM_CMEAN = 1, PARAM_ALPHA = 2 };

#ifdef NEMO_KURAMOTO_LAG
constexpr bool LAG = true;
#else
constexpr bool LAG = false;
#endif

/* 2.0f * M_PI is evaluated in double (Kuramoto.cpp:124) */
constexpr double TWO_PI = 2.0 * 3.14159265358979323846;

/* the phase stored after a step: fmodf(phase, float(2 pi)) + (2 pi if
 * phase < 0), the sum in double (Kuramoto.cpp:124, 153) */
__device__ inline float wrap_phase(float phase)
{
	const float r = nemo_math::fmodf_exact(phase, float(TWO_PI));
	return float(double(r) + (phase < 0.0f ? TWO_PI : 0.0));
}

/* One RK4 step of oscillator nl at cycle `cycle` (Kuramoto.cpp:100-126);
 * returns the new phase and sets *range when a sine argument left
 * (-120, 120).  With `levelled`, a source read at slot (cycle + 1) % H --
 * the slot this step writes -- is the source's new phase when its index is
 * below nl (the CPU backend updated it earlier in its loop) and its old
 * phase, kept in scratch, otherwise. */
__device__ inline float kuramoto_phase(const nemo_hip_neuron_group& g, const nemo_hip_step_buffers& b, unsigned nl,
		uint32_t cycle, bool levelled, bool* range)
{
	const size_t P = g.pitch;
	const unsigned H = g.history;
	const float* theta = g.state; /* slot s at theta + s * P (one state variable) */
	const uint32_t next = (cycle + 1) % H;
	const float f = g.param[PARAM_FREQUENCY * P + nl];
	const float cm = g.param[PARAM_CMEAN * P + nl];
	const float Cmean = cm > 0 ? cm : 1;
	const float alpha = LAG ? g.param[PARAM_ALPHA * P + nl] : 0.0f;
	const float h = 0.05f;
	const float target = theta[size_t(cycle % H) * P + nl];
	const unsigned ng = g.start + nl;
	const uint32_t r0 = b.rcm_all.ptr[ng], r1 = b.rcm_all.ptr[ng + 1];
	/* sumN (Kuramoto.cpp:22-34) over the synapses in reverse-matrix order */
	auto sumN = [&](float targetPhase) {
		float sum = 0.0f;
		for(uint32_t k = r0; k < r1; ++k) {
			const uint32_t sd = b.rcm_all.source_delay[k];
			const unsigned d = (sd >> 26) + 1;
			const uint32_t slot = (cycle - uint32_t(d - 1)) % H;
			const unsigned src = (sd & 0x3ffffffu) - g.start;
			const float sourcePhase = levelled && slot == next && src >= nl ? g.scratch[src]
				: theta[size_t(slot) * P + src];
			const float x = LAG ? sourcePhase - targetPhase - alpha : sourcePhase - targetPhase;
			*range = *range || __builtin_fabsf(x) >= 120.0f;
			sum += b.rcm_all.weight[k] * nemo_math::sinf_exact(x);
		}
		return sum;
	};
	const float k0 = f + (sumN(target) / Cmean);
	const float k1 = f + (sumN(target + k0 * 0.5f * h) / Cmean);
	const float k2 = f + (sumN(target + k1 * 0.5f * h) / Cmean);
	const float k3 = f + (sumN(target + k2 * h) / Cmean);
	return wrap_phase(target + h * (k0 + 2 * k1 + 2 * k2 + k3) / 6.0f);
}

__device__ inline uint32_t cycle_of(const nemo_hip_step_buffers& b)
{
	/* the reference's cycle is a 32-bit unsigned (H divides 2^32) */
	return uint32_t(nemo_dev::current_step(b) - b.clock->cycle_origin);
}

/* One launch per step: each thread one oscillator of its tile, then the
 * tile's (empty) firing record. */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
kuramoto_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned ng = g.start + nl;
	const uint64_t hist = nemo_dev::load_history(b, ng, valid);
	if(valid) {
		const uint32_t cycle = cycle_of(b);
		bool range = false;
		const float phase = kuramoto_phase(g, b, nl, cycle, false, &range);
		g.state[size_t((cycle + 1) % g.history) * g.pitch + nl] = phase;
		if(range) atomicOr(b.status, NEMO_HIP_STATUS_SINF_RANGE);
	}
	nemo_dev::record_firing(b, L, ng, valid, false, hist);
	nemo_dev::stdp_tile(b, L, ng, valid, hist);
}

/* Update passes (nemo_hip_neuron_group.levels): first the old contents of
 * the slot the step writes into scratch, then pass by pass the passes'
 * oscillators, then the tiles' firing records. */
__global__ void __launch_bounds__(NEMO_UPDATE_BLOCK)
kuramoto_keep_next(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	if(nl < g.count) g.scratch[nl] = g.state[size_t((cycle_of(b) + 1) % g.history) * g.pitch + nl];
}

__global__ void __launch_bounds__(NEMO_UPDATE_BLOCK)
kuramoto_pass(nemo_hip_neuron_group g, nemo_hip_step_buffers b, uint32_t first, uint32_t count)
{
	const unsigned i = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	if(i >= count) return;
	const unsigned nl = g.level_idx[first + i];
	const uint32_t cycle = cycle_of(b);
	bool range = false;
	const float phase = kuramoto_phase(g, b, nl, cycle, true, &range);
	g.state[size_t((cycle + 1) % g.history) * g.pitch + nl] = phase;
	if(range) atomicOr(b.status, NEMO_HIP_STATUS_SINF_RANGE);
}

template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
kuramoto_record(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned ng = g.start + nl;
	const uint64_t hist = nemo_dev::load_history(b, ng, valid);
	nemo_dev::record_firing(b, L, ng, valid, false, hist);
	nemo_dev::stdp_tile(b, L, ng, valid, hist);
}

/* Kuramoto.cpp:137-159: the history before the first cycle, each
 * oscillator run backwards one omega per slot without coupling, from slot 0
 * into slots H-1, H-2, ..., 1 */
__global__ void __launch_bounds__(NEMO_UPDATE_BLOCK)
kuramoto_init(nemo_hip_neuron_group g)
{
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	if(nl >= g.count) return;
	const size_t P = g.pitch;
	const unsigned H = g.history;
	const float f = g.param[PARAM_FREQUENCY * P + nl];
	for(unsigned t = 0; t + 1 < H; ++t) {
		const uint32_t current = 0u - t;
		const uint32_t previous = 0u - (t + 1u);
		const float phase = g.state[size_t(current % H) * P + nl] - f;
		g.state[size_t(previous % H) * P + nl] = wrap_phase(phase);
	}
}

} // namespace

extern "C" hipError_t
hip_update_neurons(hipStream_t stream, const nemo_hip_neuron_group* g, const nemo_hip_step_buffers* b)
{
	if(g->count == 0) return hipSuccess;
	if(!b->rcm_all.ptr || !b->status || g->history < 1) return hipErrorInvalidValue;
	const unsigned tiles = (g->count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK;
	/* the group's tiles, then the propagate blocks the launch carries */
	const unsigned blocks = tiles + b->co.deliver_blocks + b->co.stdp_blocks;
	const bool co = b->co.deliver_blocks || b->co.stdp_blocks;
	if(!g->level_ptr) {
		if(co) {
			hipExtLaunchKernelGGL(kuramoto_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes,
					stream, b->time_start, b->time_stop, 0, *g, *b);
		} else {
			hipExtLaunchKernelGGL(kuramoto_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes,
					stream, b->time_start, b->time_stop, 0, *g, *b);
		}
		return hipGetLastError();
	}
	if(!g->scratch || !g->level_idx || g->levels < 1) return hipErrorInvalidValue;
	hipExtLaunchKernelGGL(kuramoto_keep_next, dim3(tiles), dim3(NEMO_UPDATE_BLOCK), 0, stream, b->time_start,
			nullptr, 0, *g, *b);
	hipError_t e = hipGetLastError();
	for(unsigned l = 0; l < g->levels && e == hipSuccess; ++l) {
		const uint32_t first = g->level_ptr[l], count = g->level_ptr[l + 1] - first;
		if(count == 0) continue;
		hipLaunchKernelGGL(kuramoto_pass, dim3((count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK),
				dim3(NEMO_UPDATE_BLOCK), 0, stream, *g, *b, first, count);
		e = hipGetLastError();
	}
	if(e != hipSuccess) return e;
	if(co) {
		hipExtLaunchKernelGGL(kuramoto_record<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				nullptr, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(kuramoto_record<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				nullptr, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t stream, const nemo_hip_neuron_group* g, const nemo_hip_step_buffers*)
{
	if(g->count == 0 || g->history < 2) return hipSuccess;
	hipLaunchKernelGGL(kuramoto_init, dim3((g->count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK),
			dim3(NEMO_UPDATE_BLOCK), 0, stream, *g);
	return hipGetLastError();
}
