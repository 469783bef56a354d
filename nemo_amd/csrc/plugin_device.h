/*
 * Device helpers shared by the neuron-model plugin kernels.
 *
 * Every plugin kernel is launched with blocks of NEMO_UPDATE_BLOCK threads,
 * one neuron per thread, and every thread of a block must reach
 * nemo_record_firing() (it contains block barriers).
 */
#ifndef NEMO_AMD_PLUGIN_DEVICE_H
#define NEMO_AMD_PLUGIN_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nemo_hip_plugin.h"
#include "glibc_math.h"

#define NEMO_UPDATE_BLOCK 256

namespace nemo_dev {

/* Saturate a 64-bit fixed-point sum to the int32 range and convert to float,
 * as NeMo's wfx_toFloat (src/nemo/fixedpoint.cpp:40-49). */
__device__ inline float wfx_to_float(int64_t v, unsigned fbits)
{
	v = v > int64_t(INT32_MAX) ? int64_t(INT32_MAX) : v;
	v = v < int64_t(INT32_MIN) ? int64_t(INT32_MIN) : v;
	return float(int32_t(v)) / float(1u << fbits);
}

/* Read the excitatory and inhibitory current of neuron ng and clear the
 * accumulators for the next step (cpu/Simulation.cpp:279-288). */
__device__ inline void take_current(const nemo_hip_step_buffers& b, unsigned ng, float& E, float& I)
{
	int64_t e = b.acc[ng];
	int64_t i = b.acc[b.acc_pitch + ng];
	b.acc[ng] = 0;
	b.acc[b.acc_pitch + ng] = 0;
	E = wfx_to_float(e, b.fbits);
	I = wfx_to_float(i, b.fbits);
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

/* Shift the firing history and append fired neurons to the fired log.  One
 * atomic on the log head per block. */
__device__ inline void record_firing(const nemo_hip_step_buffers& b, unsigned ng, bool valid, bool fired)
{
	__shared__ unsigned s_count;
	__shared__ unsigned long long s_base;
	if(valid) {
		uint64_t h = b.hist[ng];
		b.hist[ng] = (h << 1) | uint64_t(fired ? 1 : 0);
	}
	if(threadIdx.x == 0) s_count = 0;
	__syncthreads();
	const bool f = valid && fired;
	const uint64_t mask = __ballot(f);
	const unsigned lane = threadIdx.x & 63;
	unsigned waveBase = 0;
	if(mask) {
		if(lane == 0) {
			waveBase = atomicAdd(&s_count, unsigned(__popcll(mask)));
		}
		waveBase = __shfl(waveBase, 0);
	}
	__syncthreads();
	if(threadIdx.x == 0 && s_count) {
		s_base = atomicAdd((unsigned long long*)&b.clock->log_head, (unsigned long long)s_count);
	}
	__syncthreads();
	if(f) {
		const unsigned rank = unsigned(__popcll(mask & ((uint64_t(1) << lane) - 1)));
		b.fired_log[(s_base + waveBase + rank) & b.log_mask] = ng;
	}
}

} // namespace nemo_dev

#endif
