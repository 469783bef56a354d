/*
 * Quadratic integrate-and-fire plugin (libQIF.so).
 *
 * Bit-exact with src/nemo/cpu/plugins/QIF.cpp:16-107: ten RK4 substeps of
 * 0.1 ms on dv/dt = a v (v - 1) + 0.015 I, skipped once v >= 1; reset v = 0.
 * Per-type arrays are indexed by the neuron's index within the type (the
 * reference indexes them by simulation index, which is only in bounds when
 * the type is the first one added; both agree in that case).
 * Parameter a and state v use the indices of src/nemo/plugins/QIF.h.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
qif_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block<true>(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned ng = g.start + nl;
	/* state, parameter and history load while the input is gathered */
	uint64_t h = 0;
	float v = 0.0f, a = 0.0f;
	if(!nemo_dev::gather_input<CO ? NEMO_COLAUNCH_BLOCK : NEMO_TILE>(b, g, L, [&]() {
		h = nemo_dev::load_history(b, ng, valid);
		v = valid ? g.state[nl] : 0.0f;
		a = valid ? g.param[nl] : 0.0f;
	}))
		return; /* helpers end here */
	bool fired = false;

	if(valid) {
		float Excit, Inhib;
		nemo_dev::take_current(b, L.input, ng, Excit, Inhib);
		const float Exter = nemo_dev::take_external(b, ng);
		const float dt = 0.1f;
		const float a1 = 0.5f * dt;
		const float a2 = dt;
		const float bb1 = dt / 6.0f;
		const unsigned inc_max = (unsigned)(1 / dt);
		for(unsigned t = 0; t < inc_max; ++t) {
			if(!fired) {
				float I = Excit + Inhib + Exter;
				I = I * 0.015f;
				float k1 = a * v * (v - 1) + I;
				float x2 = v + a1 * k1;
				float k2 = a * x2 * (x2 - 1) + I;
				float x3 = v + a1 * k2;
				float k3 = a * x3 * (x3 - 1) + I;
				float x4 = v + a2 * k3;
				float k4 = a * x4 * (x4 - 1) + I;
				v = v + bb1 * (k1 + 2 * k2 + 2 * k3 + k4);
				fired = v >= 1;
			}
		}
		fired = fired || nemo_dev::stimulated(b, ng);
		if(fired) {
			v = 0;
		}
		g.state[nl] = v;
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
		hipExtLaunchKernelGGL(qif_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(qif_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess;
}
