/*
 * The CPU-only models of NeMo (src/nemo/cpu/plugins, cuda=false in their
 * .ini files), one plugin each (plugins/<Model>.hip defines NEMO_VARIANT and
 * includes this file).  Every update is the reference's arithmetic in its
 * order, FP contraction off: literals written as doubles in the reference
 * make the same expressions double here.  Per-type arrays are indexed by the
 * neuron's index within its type (the reference's plugins index them with
 * that local index too, except where noted).
 *
 *   VARIANT_CHAOS     IzhikevichChaos.cpp:14-101
 *   VARIANT_SYN       IzhikevichSyn.cpp:14-140 (plugins/IzhikevichSyn.h)
 *   VARIANT_2007      Izhikevich2007.cpp:14-134 (plugins/Izhikevich2007.h)
 *   VARIANT_DOPSYN    Izhikevich2007DopSyn.cpp:14-156
 *   VARIANT_BAYESIAN  BayesianNeuron.cpp:14-93 (plugins/BayesianNeuron.h)
 *   VARIANT_INHOMOGENEOUS_POISSON  InhomogeneousPoissonSource.cpp:6-27
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

#define VARIANT_CHAOS 1
#define VARIANT_SYN 2
#define VARIANT_2007 3
#define VARIANT_DOPSYN 4
#define VARIANT_BAYESIAN 5
#define VARIANT_INHOMOGENEOUS_POISSON 6

namespace {

/* plugins/Izhikevich2007DopSyn.h (double literals) */
constexpr double MSN_BETA1 = 6.3, MSN_BETA2 = 0.215, MSN_K = 0.0289, MSN_L = 0.331, MSN_ALPHA = 0.032;

/* One step of neuron nl (valid); returns whether it fired. */
__device__ inline bool variant_step(const nemo_hip_neuron_group& g, const nemo_hip_step_buffers& b, unsigned nl,
		unsigned ng, float E, float In)
{
	const size_t P = g.pitch;
	auto par = [&](unsigned i) { return g.param[i * P + nl]; };
	bool fired = false;
#if NEMO_VARIANT == VARIANT_INHOMOGENEOUS_POISSON
	/* the rate (Hz) is the held external current, not cleared */
	const float rate = b.ext_held[ng];
	const unsigned p0 = unsigned(0.001 * rate * float(1 << 16));
	uint4 s = reinterpret_cast<const uint4*>(g.rng)[nl];
	const unsigned p1 = nemo_dev::urand(s) & 0xffff;
	reinterpret_cast<uint4*>(g.rng)[nl] = s;
	fired = p1 < p0;
	(void) E;
	(void) In;
#elif NEMO_VARIANT == VARIANT_BAYESIAN
	float* L0 = g.state;
	float* G0 = g.state + P;
	float I = E + In + nemo_dev::take_external(b, ng);
	float L = L0[nl];
	float G = G0[nl];
	const float rOn = par(0), rOff = par(1), g0 = par(2);
	for(int t = 0; t < 4; ++t) {
		if(!fired) {
			L += 0.25f * (rOn * (1 + nemo_math::expf_exact(-L)) - rOff * (1 + nemo_math::expf_exact(L)) + I);
			G += 0.25f * (rOn * (1 + nemo_math::expf_exact(-G)) - rOff * (1 + nemo_math::expf_exact(G)));
			fired = double(L) > (double(G) + double(g0) / 2.0);
		}
	}
	fired = fired || nemo_dev::stimulated(b, ng);
	if(fired) G += g0;
	L0[nl] = L;
	G0[nl] = G;
	return fired;
#else
	float* u0 = g.state;
	float* v0 = g.state + P;
	float u = u0[nl];
	float v = v0[nl];
#if NEMO_VARIANT == VARIANT_CHAOS
	float I = E + In + nemo_dev::take_external(b, ng);
	uint4 s = reinterpret_cast<const uint4*>(g.rng)[nl];
	const float randNoise = par(4) * nemo_dev::nrand(s);
	reinterpret_cast<uint4*>(g.rng)[nl] = s;
	const float a = par(0), bb = par(1);
	for(int t = 0; t < 20; ++t) {
		if(!fired) {
			v = float(double(v) + double(0.05f) * ((((0.04 * double(v) + 5.0) * double(v)) + 140.0) - double(u)
						+ double(I)));
			u = u + 0.05f * (a * (bb * v - u) + randNoise);
			fired = v >= 30.0f;
		}
	}
	const float c = par(2), d = par(3);
#elif NEMO_VARIANT == VARIANT_SYN
	float* ge0 = g.state + 2 * P;
	float* gi0 = g.state + 3 * P;
	float ge = ge0[nl] * par(6);
	float gi = gi0[nl] * par(9);
	ge += E;
	gi += In;
	if(ge > par(7)) ge = par(7);
	if(gi < -par(10)) gi = -par(10);
	ge0[nl] = ge;
	gi0[nl] = gi;
	float I = ge * (par(8) - v) - gi * (par(11) - v) + nemo_dev::take_external(b, ng);
	const float sigma = par(5);
	if(sigma != 0.0f) {
		uint4 s = reinterpret_cast<const uint4*>(g.rng)[nl];
		I += sigma * nemo_dev::nrand(s);
		reinterpret_cast<uint4*>(g.rng)[nl] = s;
	}
	const float a = par(0), bb = par(1), vpeak = par(4);
	for(int t = 0; t < 4; ++t) {
		if(!fired) {
			v = float(double(v) + double(0.25f) * ((((0.04 * double(v) + 5.0) * double(v)) + 140.0) - double(u)
						+ double(I)));
			u = u + 0.25f * (a * (bb * v - u));
			fired = v >= vpeak;
		}
	}
	const float c = par(2), d = par(3);
#else /* VARIANT_2007, VARIANT_DOPSYN */
	constexpr bool DOP = NEMO_VARIANT == VARIANT_DOPSYN;
	constexpr unsigned OFF = DOP ? 2 : 0; /* d1, d2 sit between sigma and the synapse parameters */
	float* ge0 = g.state + 2 * P;
	float* gi0 = g.state + 3 * P;
	const float a = par(0), bb = par(1), vpeak = par(4), Vr = par(5), Vt = par(6), k = par(7), C = par(8);
	const float sigma = par(9);
	float ge = ge0[nl] * par(10 + OFF);
	float gi = gi0[nl] * par(13 + OFF);
	float I;
	float d1 = 0.0f, d2 = 0.0f;
	if(DOP) {
		d1 = par(10);
		d2 = par(11);
		ge = float(double(ge) + (double(1.0f) + MSN_BETA1 * double(d1)) * (double(1.0f) - MSN_BETA2 * double(d2))
				* double(E));
		gi += In;
		if(ge > par(11 + OFF)) ge = par(11 + OFF);
		if(gi < -par(14 + OFF)) gi = -par(14 + OFF);
		I = ge * (par(12 + OFF) - v) - gi * (par(15 + OFF) - v) + nemo_dev::take_external(b, ng);
	} else {
		if(E > 0.0f) ge = 1.0f;
		if(In > 0.0f) gi = 1.0f;
		I = par(11) * ge * (par(12) - v) + par(14) * gi * (par(15) - v) + nemo_dev::take_external(b, ng);
	}
	ge0[nl] = ge;
	gi0[nl] = gi;
	if(sigma != 0.0f) {
		uint4 s = reinterpret_cast<const uint4*>(g.rng)[nl];
		I += DOP ? sigma * nemo_dev::nrand(s) : C * sigma * nemo_dev::nrand(s);
		reinterpret_cast<uint4*>(g.rng)[nl] = s;
	}
	const float oneOverC = 1.0f / C;
	float d = par(3);
	for(int t = 0; t < 4; ++t) {
		if(!fired) {
			if(DOP) {
				const float msn_Vr = float(double(Vr) * (1 + MSN_K * double(d1)));
				d = float(double(par(3)) * (1 - MSN_L * double(d1)));
				const float msn_k = float(double(k) * (1 - MSN_ALPHA * double(d2)));
				v += 0.25f * oneOverC * (msn_k * (v - msn_Vr) * (v - Vt) - u + I);
				u += 0.25f * (a * (bb * (v - msn_Vr) - u));
			} else {
				v += 0.25f * oneOverC * (k * (v - Vr) * (v - Vt) - u + I);
				u += 0.25f * (a * (bb * (v - Vr) - u));
			}
			fired = v >= vpeak;
		}
	}
	const float c = par(2);
#endif
	fired = fired || nemo_dev::stimulated(b, ng);
	if(fired) {
		v = c;
		u += d;
	}
	u0[nl] = u;
	v0[nl] = v;
	return fired;
#endif
	return fired || nemo_dev::stimulated(b, ng);
}

template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
variant_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
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
		nemo_dev::take_current(b, L.input, ng, E, I);
		fired = variant_step(g, b, nl, ng, E, I);
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
		hipExtLaunchKernelGGL(variant_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(variant_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, b->time_stop, 0, *g, *b);
	}
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess; /* src/nemo/cpu/plugins/default_init.c:3-21: nothing to do */
}
