/*
 * Hodgkin-Huxley neuron plugin (libHH.so).
 *
 * Bit-exact with src/nemo/cpu/plugins/HH.cpp:14-138: 999 Euler substeps
 * ((int)(1/0.001f)) of the classic HH equations with glibc-exact expf; the
 * neuron fires on the first downturn of v (change < -1e-9) once the timer
 * cycle is >= 10.  State v, n, m, h, dir (HH.ini: 5 state variables, no
 * parameters).  Per-type arrays are indexed within the type (see QIF).
 *
 * A step is three launches when the group has its four scratch planes
 * (HH.ini [hip] scratch-planes=4): the tiles' input gather (one thread per
 * neuron, as every model), the substeps with sixteen lanes per neuron in
 * one-wave blocks -- lane q evaluates the q-th of the six rate functions
 * (each an expf), the lanes exchange them and all carry the same state -- and
 * the tiles' firing record.  One wave of a tile block otherwise spends each
 * substep on six exponentials for its 64 neurons, one wave per SIMD on a few
 * CUs; the split spreads the same arithmetic, operation for operation, over
 * sixteen times as many waves.  Without scratch planes the single launch runs.
 */
#include <hip/hip_ext.h>

#include "../plugin_device.h"

namespace {

using nemo_math::expf_exact;

/* HH.cpp:26-39 */
constexpr float DT = 0.001f;
constexpr float G_NA = 120.0f;
constexpr float G_K = 36.0f;
constexpr float G_L = 0.3f;
constexpr float E_NA = 115.0f - 65.0f;
constexpr float E_K = -12.0f - 65.0f;
constexpr float E_L = 10.6f - 65.0f;
constexpr float C_M = 1.0f;
constexpr float REV_E = 0.0f;
constexpr float REV_I = -70.0f;

/* scratch planes of the split step */
enum { SCR_E = 0, SCR_I = 1, SCR_X = 2, SCR_FIRED = 3 };
constexpr unsigned LANES = 16;     /* per neuron in the substep launch: one DPP row */
constexpr unsigned SUB_BLOCK = 64; /* one wave: eight neurons */

/* The value of lane Q of this lane's row of 16 (DPP row_share). */
template<int Q>
__device__ inline float row_lane(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + Q, 0xf, 0xf, false));
}

/* The gating and membrane update of one substep from the six rates (HH.cpp:44-62). */
__device__ inline void hh_substep(float& v, float& n, float& m, float& h, float& dir, bool& fired, uint64_t cycle,
		float Excit, float Inhib, float Exter, float alphan, float alpham, float alphah, float betan,
		float betam, float betah)
{
	const float I = (Excit * (REV_E - v)) + (Inhib * ((REV_I - v) / -1)) + Exter;
	m = m + DT * (alpham * (1.0f - m) - betam * m);
	n = n + DT * (alphan * (1.0f - n) - betan * n);
	h = h + DT * (alphah * (1.0f - h) - betah * h);
	const float Ik = G_NA * (m * m * m) * h * (v - E_NA) + G_K * (n * n * n * n) * (v - E_K) + G_L * (v - E_L);
	const float newv = v + DT * (-Ik + I) / C_M;
	const float new_dir = (newv - v);
	const float change = (dir < 0) | (newv < -45) ? 0 : new_dir;
	dir = new_dir;
	if(!fired && cycle >= 10) fired = double(change) < -0.000000001;
	v = newv;
}

/* Single launch: gather, substeps and firing record per tile. */
/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
hh_update(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block<true>(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	if(!nemo_dev::gather_input<CO ? NEMO_COLAUNCH_BLOCK : NEMO_TILE>(b, g, L)) return; /* helpers end here */
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned nn = g.start + nl;
	const uint64_t hist = nemo_dev::load_history(b, nn, valid);
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
		nemo_dev::take_current(b, L.input, nn, Excit, Inhib);
		const float Exter = nemo_dev::take_external(b, nn);
		const uint64_t cycle = nemo_dev::current_step(b) - b.clock->cycle_origin;
		const int inc_max = (int)(1 / DT);
		for(int t = 0; t < inc_max; ++t) {
			const float alphan = (0.1f - 0.01f * (v + 65.0f)) / (expf_exact(1.0f - 0.1f * (v + 65.0f)) - 1.0f);
			const float alpham = (2.5f - 0.1f * (v + 65.0f)) / (expf_exact(2.5f - 0.1f * (v + 65.0f)) - 1.0f);
			const float alphah = 0.07f * expf_exact(-(v + 65.0f) / 20.0f);
			const float betan = 0.125f * expf_exact(-(v + 65.0f) / 80.0f);
			const float betam = 4.0f * expf_exact(-(v + 65.0f) / 18.0f);
			const float betah = 1.0f / (expf_exact(3.0f - 0.1f * (v + 65.0f)) + 1.0f);
			hh_substep(v, n, m, h, dir, fired, cycle, Excit, Inhib, Exter, alphan, alpham, alphah, betan, betam, betah);
		}
		fired = fired || nemo_dev::stimulated(b, nn);
		s[0 * P + nl] = v;
		s[1 * P + nl] = n;
		s[2 * P + nl] = m;
		s[3 * P + nl] = h;
		s[4 * P + nl] = dir;
	}
	nemo_dev::record_firing(b, L, nn, valid, fired, hist);
	nemo_dev::stdp_tile(b, L, nn, valid, hist);
}

/* Split step, launch 1: the tiles' synaptic and external input. */
/* CO: the launch may carry propagate blocks after the tiles (compiled
 * out otherwise, which leaves the update far fewer registers) */
template<bool CO>
__global__ void NEMO_UPDATE_BOUNDS(CO)
hh_input(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	if(CO && nemo_dev::colaunched_block<true>(b, g)) return;
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<CO>();
	if(!nemo_dev::gather_input<CO ? NEMO_COLAUNCH_BLOCK : NEMO_TILE>(b, g, L)) return; /* helpers end here */
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	if(nl < g.count) {
		const unsigned nn = g.start + nl;
		const size_t P = g.pitch;
		float Excit, Inhib;
		nemo_dev::take_current(b, L.input, nn, Excit, Inhib);
		g.scratch[SCR_E * P + nl] = Excit;
		g.scratch[SCR_I * P + nl] = Inhib;
		g.scratch[SCR_X * P + nl] = nemo_dev::take_external(b, nn);
	}
}

/* Split step, launch 2: the substeps, LANES lanes per neuron.  Lane q < 6
 * evaluates rate q (alphan, alpham, alphah, betan, betam, betah; lanes 6 to
 * 15 repeat betah) from exactly the expression of the single launch, and a
 * DPP row broadcast hands each rate to the whole row (the neuron); every
 * lane then applies the same update, so all lanes of a neuron hold the same
 * state and the first writes it.
 *
 * The substeps are one dependent chain, so their latency is the launch's
 * length: every lane runs the same instructions (the six rate formulas are
 * one division of selected operands: a division by 1 is exact, so
 * "fac * e / 1" and "(-x / 1)" are fac * e and -x bit for bit), the
 * exponent's division by its constant is a reciprocal product with one FMA
 * correction (nemo_math::div_by_const: the same bits, or for tiny arguments
 * the same expf), expf's special cases are selected rather than branched,
 * and its table is read from LDS. */
__global__ void __launch_bounds__(SUB_BLOCK)
hh_substeps(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	__shared__ uint64_t s_exp[32];
	if(threadIdx.x < 32) s_exp[threadIdx.x] = nemo_math::expf_table(threadIdx.x);
	__syncthreads();
	const unsigned t = blockIdx.x * SUB_BLOCK + threadIdx.x;
	const unsigned nl = t / LANES;
	const unsigned q = t % LANES;
	if(nl >= g.count) return; /* whole groups of LANES lanes */
	const size_t P = g.pitch;
	float* s = g.state;
	float v = s[0 * P + nl];
	float n = s[1 * P + nl];
	float m = s[2 * P + nl];
	float h = s[3 * P + nl];
	float dir = s[4 * P + nl];
	const float Excit = g.scratch[SCR_E * P + nl];
	const float Inhib = g.scratch[SCR_I * P + nl];
	const float Exter = g.scratch[SCR_X * P + nl];
	const uint64_t cycle = nemo_dev::current_step(b) - b.clock->cycle_origin;
	/* per lane (HH.cpp:92-98), with x = v + 65:
	 *   exponent  (lin0 - 0.1 x) / 1   (alphan 1, alpham 2.5, betah 3)
	 *             (-x) / div           (alphah 20, betan 80, betam 18)
	 *   rate      (c0 - c1 x) / (e - 1)  (alphan 0.1, 0.01; alpham 2.5, 0.1)
	 *             1 / (e + 1)            (betah)
	 *             (fac e) / 1            (alphah 0.07, betan 0.125, betam 4) */
	const unsigned r = q < 6 ? q : 5;
	const bool linear = r == 0 || r == 1 || r == 5;
	const float lin0 = r == 0 ? 1.0f : r == 1 ? 2.5f : 3.0f;
	const float adiv = linear ? 1.0f : r == 2 ? 20.0f : r == 3 ? 80.0f : 18.0f;
	const float ainv = 1.0f / adiv;
	const float fac = r == 2 ? 0.07f : r == 3 ? 0.125f : 4.0f;
	const float c0 = r == 0 ? 0.1f : 2.5f;
	const float c1 = r == 0 ? 0.01f : 0.1f;
	const float eoff = r <= 1 ? -1.0f : 1.0f;
	bool fired = false;
	const int inc_max = (int)(1 / DT);
	for(int it = 0; it < inc_max; ++it) {
		const float x = v + 65.0f;
		const float anum = linear ? lin0 - 0.1f * x : -x;
		/* anum / adiv, the same bits or (tiny anum) the same expf */
		const float e = nemo_math::expf_exact_tab(nemo_math::div_by_const(anum, adiv, ainv), s_exp);
		const float poly = c0 - c1 * x;
		const float fe = fac * e;
		const float rnum = r <= 1 ? poly : r == 5 ? 1.0f : fe;
		const float eden = e + eoff;
		const float rden = r <= 1 || r == 5 ? eden : 1.0f;
		const float rate = rnum / rden;
		const float alphan = row_lane<0>(rate);
		const float alpham = row_lane<1>(rate);
		const float alphah = row_lane<2>(rate);
		const float betan = row_lane<3>(rate);
		const float betam = row_lane<4>(rate);
		const float betah = row_lane<5>(rate);
		hh_substep(v, n, m, h, dir, fired, cycle, Excit, Inhib, Exter, alphan, alpham, alphah, betan, betam, betah);
	}
	if(q == 0) {
		s[0 * P + nl] = v;
		s[1 * P + nl] = n;
		s[2 * P + nl] = m;
		s[3 * P + nl] = h;
		s[4 * P + nl] = dir;
		g.scratch[SCR_FIRED * P + nl] = fired ? 1.0f : 0.0f;
	}
}

/* Split step, launch 3: the tiles' firing record (and per-tile STDP). */
__global__ void __launch_bounds__(NEMO_UPDATE_BLOCK)
hh_record(nemo_hip_neuron_group g, nemo_hip_step_buffers b)
{
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<false>();
	const unsigned nl = blockIdx.x * NEMO_UPDATE_BLOCK + threadIdx.x;
	const bool valid = nl < g.count;
	const unsigned nn = g.start + nl;
	const uint64_t hist = nemo_dev::load_history(b, nn, valid);
	L.push_rows[threadIdx.x] = nemo_dev::push_row(b, nn);
	bool fired = false;
	if(valid) fired = g.scratch[SCR_FIRED * g.pitch + nl] != 0.0f || nemo_dev::stimulated(b, nn);
	nemo_dev::record_firing(b, L, nn, valid, fired, hist);
	nemo_dev::stdp_tile(b, L, nn, valid, hist);
}

} // namespace

extern "C" hipError_t
hip_update_neurons(hipStream_t stream, const nemo_hip_neuron_group* g, const nemo_hip_step_buffers* b)
{
	if(g->count == 0) return hipSuccess;
	const unsigned tiles = (g->count + NEMO_UPDATE_BLOCK - 1) / NEMO_UPDATE_BLOCK;
	/* the group's tiles, then the propagate blocks the launch carries */
	const unsigned blocks = tiles + b->co.deliver_blocks + b->co.stdp_blocks;
	if(!g->scratch) {
		if(b->co.deliver_blocks || b->co.stdp_blocks) {
			hipExtLaunchKernelGGL(hh_update<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
					b->time_start, b->time_stop, 0, *g, *b);
		} else {
			hipExtLaunchKernelGGL(hh_update<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
					b->time_start, b->time_stop, 0, *g, *b);
		}
		return hipGetLastError();
	}
	if(b->co.deliver_blocks || b->co.stdp_blocks) {
		hipExtLaunchKernelGGL(hh_input<true>, dim3(blocks), dim3(NEMO_COLAUNCH_BLOCK), b->co.lds_bytes, stream,
				b->time_start, nullptr, 0, *g, *b);
	} else {
		hipExtLaunchKernelGGL(hh_input<false>, dim3(blocks), dim3(NEMO_UPDATE_BLOCK), b->co.lds_bytes, stream,
				b->time_start, nullptr, 0, *g, *b);
	}
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return e;
	const unsigned sub = unsigned((size_t(g->count) * LANES + SUB_BLOCK - 1) / SUB_BLOCK);
	hipLaunchKernelGGL(hh_substeps, dim3(sub), dim3(SUB_BLOCK), 0, stream, *g, *b);
	e = hipGetLastError();
	if(e != hipSuccess) return e;
	hipExtLaunchKernelGGL(hh_record, dim3(tiles), dim3(NEMO_UPDATE_BLOCK), 0, stream, nullptr, b->time_stop, 0, *g,
			*b);
	return hipGetLastError();
}

extern "C" hipError_t
hip_init_neurons(hipStream_t, const nemo_hip_neuron_group*, const nemo_hip_step_buffers*)
{
	return hipSuccess; /* src/nemo/cpu/plugins/default_init.c:3-21: nothing to do */
}
