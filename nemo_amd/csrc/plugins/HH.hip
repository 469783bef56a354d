/*
 * Hodgkin-Huxley neuron plugin (libHH.so).
 *
 * Bit-exact with src/nemo/cpu/plugins/HH.cpp:14-138: 999 Euler substeps
 * ((int)(1/0.001f)) of the classic HH equations with glibc-exact expf; the
 * neuron fires on the first downturn of v (change < -1e-9) once the timer
 * cycle is >= 10.  State v, n, m, h, dir (HH.ini: 5 state variables, no
 * parameters).  Per-type arrays are indexed within the type (see QIF).
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

__global__ void __launch_bounds__(NEMO_UPDATE_BLOCK)
hh_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(nemo_dev::colaunched_block(b, g)) return;
	using nemo_math::expf_exact;
	__shared__ nemo_dev::TileInput input;
	nemo_dev::gather_input(b, g, input);
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned nn = g.start + nl;
	const uint64_t h = nemo_dev::load_history(b, nn, valid);
	bool fired = false;

	if(valid) {
		const size_t P = g.pitch;
		float* s = g.state;
		float v = s[0 * P + nl];
		float n = s[1 * P + nl];
		float m = s[2 * P + nl];
		float h = s[3 * P + nl];
		float dir = s[4 * P + nl];
		float Excit, Inhib;
		nemo_dev::take_current(b, input, nn, Excit, Inhib);
		const float Exter = nemo_dev::take_external(b, nn);
		const uint64_t cycle = nemo_dev::current_step(b) - b.clock->cycle_origin;

		const float dt = 0.001f;
		const float gNa = 120.0f;
		const float gK = 36.0f;
		const float gL = 0.3f;
		const float ENa = 115.0f - 65.0f;
		const float EK = -12.0f - 65.0f;
		const float EL = 10.6f - 65.0f;
		const float C = 1.0f;
		const float RevE = 0.0f;
		const float RevI = -70.0f;
		const int inc_max = (int)(1 / dt);

		for(int t = 0; t < inc_max; ++t) {
			float I = (Excit * (RevE - v)) + (Inhib * ((RevI - v) / -1)) + Exter;
			float alphan = (0.1f - 0.01f * (v + 65.0f)) / (expf_exact(1.0f - 0.1f * (v + 65.0f)) - 1.0f);
			float alpham = (2.5f - 0.1f * (v + 65.0f)) / (expf_exact(2.5f - 0.1f * (v + 65.0f)) - 1.0f);
			float alphah = 0.07f * expf_exact(-(v + 65.0f) / 20.0f);
			float betan = 0.125f * expf_exact(-(v + 65.0f) / 80.0f);
			float betam = 4.0f * expf_exact(-(v + 65.0f) / 18.0f);
			float betah = 1.0f / (expf_exact(3.0f - 0.1f * (v + 65.0f)) + 1.0f);
			m = m + dt * (alpham * (1.0f - m) - betam * m);
			n = n + dt * (alphan * (1.0f - n) - betan * n);
			h = h + dt * (alphah * (1.0f - h) - betah * h);
			float Ik = gNa * (m * m * m) * h * (v - ENa) + gK * (n * n * n * n) * (v - EK) + gL * (v - EL);
			float newv = v + dt * (-Ik + I) / C;
			float new_dir = (newv - v);
			float change = (dir < 0) | (newv < -45) ? 0 : new_dir;
			dir = new_dir;
			if(!fired && cycle >= 10)
				fired = double(change) < -0.000000001;
			v = newv;
		}
		fired = fired || nemo_dev::stimulated(b, nn);
		s[0 * P + nl] = v;
		s[1 * P + nl] = n;
		s[2 * P + nl] = m;
		s[3 * P + nl] = h;
		s[4 * P + nl] = dir;
	}
	nemo_dev::record_firing(b, nn, valid, fired, h);
	nemo_dev::stdp_tile(b, nn, valid, h);
}

} // namespace

extern "C" hipError_t
hip_update_neurons(hipStream_t stream, const nemo_hip_neuron_group* g, const nemo_hip_step_buffers* b)
{
	if(g->count == 0) return hipSuccess;
	/* the group's tiles, then the propagate blocks the launch carries */
	const unsigned blocks = (g->count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK + b->co.deliver_blocks
		+ b->co.stdp_blocks;
	hipExtLaunchKernelGGL(hh_update, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream, b->time_start,
			b->time_stop, 0, *g, *b);
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess;
}
