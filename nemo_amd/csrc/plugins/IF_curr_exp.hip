/*
 * Leaky integrate-and-fire neuron with exponentially decaying synaptic
 * current (libIF_curr_exp.so).
 *
 * Bit-exact with src/nemo/cpu/plugins/IF_curr_exp.cpp:27-122:
 *   Ie = (1 - 1/tau_syn_E) Ie + E,  Ii = (1 - 1/tau_syn_I) Ii + I
 *   outside the refractory period (lastfired > tau_refrac):
 *     v += (Ie + Ii + ext + i_offset) / c_m + (v_rest - v) / tau_m
 *   fire when not refractory and v > v_thresh (the CPU backend's strict
 *   test), or when stimulated; then lastfired = 1, v = v_reset, else
 *   lastfired += 1.
 * Divisions are correctly rounded (-fhip-fp32-correctly-rounded-divide-sqrt).
 * Indices from src/nemo/plugins/IF_curr_exp.h:4-18.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

enum { V_REST = 0, V_RESET = 1, C_M = 2, TAU_M = 3, TAU_REFRAC = 4, TAU_SYN_E = 5, TAU_SYN_I = 6,
	V_THRESH = 7, I_OFFSET = 8 };
enum { S_V = 0, S_IE = 1, S_II = 2, S_LASTFIRED = 3 };

/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
if_curr_exp_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block<true>(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	if(!nemo_dev::gather_input<CO ? NEMO_COLAUNCH_BLOCK : NEMO_TILE>(b, g, L)) return; /* helpers end here */
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned ng = g.start + nl;
	const uint64_t h = nemo_dev::load_history(b, ng, valid);
	bool fired = false;

	if(valid) {
		const size_t P = g.pitch;
		const float* p = g.param;
		float* s = g.state;
		float E, In;
		nemo_dev::take_current(b, L.input, ng, E, In);
		float Ie = ((1.0f - 1.0f / p[TAU_SYN_E * P + nl]) * s[S_IE * P + nl]) + E;
		float Ii = ((1.0f - 1.0f / p[TAU_SYN_I * P + nl]) * s[S_II * P + nl]) + In;
		float I = Ie + Ii + nemo_dev::take_external(b, ng) + p[I_OFFSET * P + nl];
		s[S_IE * P + nl] = Ie;
		s[S_II * P + nl] = Ii;

		float v = s[S_V * P + nl];
		const float lastfired = s[S_LASTFIRED * P + nl];
		const bool refractory = lastfired <= p[TAU_REFRAC * P + nl];
		if(!refractory) {
			const float c_m = p[C_M * P + nl];
			const float v_rest = p[V_REST * P + nl];
			const float tau_m = p[TAU_M * P + nl];
			v += I / c_m + (v_rest - v) / tau_m;
		}
		fired = (!refractory && v > p[V_THRESH * P + nl]) || nemo_dev::stimulated(b, ng);
		if(fired) {
			s[S_LASTFIRED * P + nl] = 1;
			s[S_V * P + nl] = p[V_RESET * P + nl];
		} else {
			s[S_LASTFIRED * P + nl] = lastfired + 1;
			s[S_V * P + nl] = v;
		}
	}
	nemo_dev::record_firing(b, L, ng, valid, fired, h);
	nemo_dev::stdp_tile(b, L, ng, valid, h);
}

} // namespace

extern "C" hipError_t
hip_update_neurons(hipStream_t stream, const nemo_hip_neuron_group* g, const nemo_hip_step_buffers* b)
{
	if(g->count == 0) return hipSuccess;
	/* the group's tiles, then the propagate blocks the launch carries */
	const unsigned blocks = (g->count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK + b->co.deliver_blocks
		+ b->co.stdp_blocks;
	if(b->co.deliver_blocks || b->co.stdp_blocks) {
		hipExtLaunchKernelGGL(if_curr_exp_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(if_curr_exp_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess;
}
