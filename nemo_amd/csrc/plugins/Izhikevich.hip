/*
 * Izhikevich neuron model plugin (libIzhikevich.so).
 *
 * Bit-exact with the CPU model src/nemo/cpu/plugins/Izhikevich.cpp:14-98:
 *   I = (E + I) + ext;  I += sigma * nrand  (when sigma != 0)
 *   4 substeps of 0.25 ms, each skipped once the neuron has fired:
 *     v = float(double(v) + 0.25 * (((0.04 v + 5) v + 140) - u + I))   [FP64]
 *     u = u + 0.25f * (a * (b v - u))                                  [FP32]
 *     fired = v >= 30
 *   fired |= fstim; on firing v = c, u += d.
 * The reference writes 0.04, 5.0 and 140.0 as double literals, so its v update
 * is evaluated in double; this kernel does the same, with FP contraction off
 * (the build passes -ffp-contract=off).
 * Parameters a, b, c, d, sigma and state u, v use the indices of
 * src/nemo/plugins/Izhikevich.h:6-15.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

enum { PARAM_A = 0, PARAM_B = 1, PARAM_C = 2, PARAM_D = 3, PARAM_SIGMA = 4 };
enum { STATE_U = 0, STATE_V = 1 };

/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
izhikevich_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block<true>(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned ng = g.start + nl;
	const size_t P = g.pitch;

	/* the neuron's parameters, state, RNG and history are requested right
	 * after the tile's input table, so that their latency overlaps the gather */
	float sigma = 0.0f, a = 0.0f, bb = 0.0f, c = 0.0f, d = 0.0f, u = 0.0f, v = 0.0f;
	uint4 s = make_uint4(0, 0, 0, 0);
	uint64_t h = 0;
	/* in a carrying launch the block's other waves help with the gather */
	if(!nemo_dev::gather_input<CO ? NEMO_COLAUNCH_BLOCK : NEMO_TILE>(b, g, L, [&]() {
		if(valid) {
			sigma = g.param[PARAM_SIGMA * P + nl];
			a = g.param[PARAM_A * P + nl];
			bb = g.param[PARAM_B * P + nl];
			c = g.param[PARAM_C * P + nl];
			d = g.param[PARAM_D * P + nl];
			u = g.state[STATE_U * P + nl];
			v = g.state[STATE_V * P + nl];
			if(g.rng) s = reinterpret_cast<const uint4*>(g.rng)[nl];
			h = nemo_dev::load_history(b, ng, true);
		}
	}))
		return;
	bool fired = false;

	if(valid) {
		float E, In;
		nemo_dev::take_current(b, L.input, ng, E, In);
		float I = E + In + nemo_dev::take_external(b, ng);

		if(sigma != 0.0f) {
			I += sigma * nemo_dev::nrand(s);
			reinterpret_cast<uint4*>(g.rng)[nl] = s;
		}

#pragma unroll
		for(int t = 0; t < 4; ++t) {
			if(!fired) {
				const double vd = double(v);
				v = float(vd + double(0.25f) * ((((0.04 * vd + 5.0) * vd) + 140.0) - double(u) + double(I)));
				u = u + 0.25f * (a * (bb * v - u));
				fired = v >= 30.0f;
			}
		}

		fired = fired || nemo_dev::stimulated(b, ng);
		if(fired) {
			v = c;
			u += d;
		}
		g.state[STATE_U * P + nl] = u;
		g.state[STATE_V * P + nl] = v;
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
		hipExtLaunchKernelGGL(izhikevich_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(izhikevich_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess; /* src/nemo/cpu/plugins/default_init.c:3-21: nothing to do */
}
