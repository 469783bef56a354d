/*
 * Poisson spike source plugin (libPoissonSource.so).
 *
 * Bit-exact with src/nemo/cpu/plugins/PoissonSource.cpp:6-30: the neuron
 * fires when (urand & 0xffff) < unsigned(rate * 2^16), or when stimulated.
 * One xorshift draw per neuron per step.  Parameter 0 is the rate per ms.
 * Synaptic and external currents are ignored (and, as in the reference, the
 * external current is not cleared); the accumulators are still read and cleared.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
poisson_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
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
		float E, I;
		nemo_dev::take_current(b, L.input, ng, E, I); /* clears the overflow accumulators */
		const unsigned p0 = unsigned(g.param[nl] * float(1 << 16));
		uint4* rng = reinterpret_cast<uint4*>(g.rng);
		uint4 s = rng[nl];
		const unsigned p1 = nemo_dev::urand(s) & 0xffff;
		rng[nl] = s;
		fired = p1 < p0 || nemo_dev::stimulated(b, ng);
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
		hipExtLaunchKernelGGL(poisson_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(poisson_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess;
}
