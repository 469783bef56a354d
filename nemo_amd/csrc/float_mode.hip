/*
 * Float-weight mode and its adjoint: the Izhikevich network with FP32
 * weights and FP32 current sums, differentiable through a surrogate of the
 * spike threshold.  No counterpart in the reference (its plain mode is the
 * fixed-point one, src/nemo/cpu/Simulation.cpp:259-308 and
 * src/nemo/cpu/plugins/Izhikevich.cpp:14-98, whose neuron step this mode
 * keeps).
 *
 * Model, per 1 ms step t and neuron j (weights in float; their sum, the
 * state v, u and all adjoint quantities in double):
 *   I    = x[t][j] + sum over synapses (i -> j, delay d, weight w) of w * S[t-d][i]
 *   4 substeps k (h = 0.25), each only while the neuron has not crossed 30 mV
 *   in an earlier substep (mask m_k):
 *     v <- v + h (0.04 v^2 + 5 v + 140 - u + I)
 *     u <- u + h a (b v - u)                                (new v)
 *   S[t][j] = H(v - 30);  v <- v + S (c - v);  u <- u + S d;  V[t][j] = v
 * Backward: reverse-mode through the same expression, with the masks m_k and
 * the spike value in the reset held constant (the reset is not
 * differentiated through the threshold), and dH/dv replaced by the
 * fast-sigmoid surrogate 1 / (1 + 10 |v - 30|)^2 where S feeds the synapses
 * and the loss.
 *
 * Layout: neurons 0..N-1; the forward matrix holds each source's synapses
 * (all delays) contiguously, fcm_td = target | (delay-1) << 26; the reverse
 * matrix holds each target's incoming synapses, rcm_sd = source | (delay-1)
 * << 26.  Weights are gathered per call into both orders.  S, V, x and the
 * gradients are [T][N] row-major float.
 *
 * One wave (64 lanes) per neuron: lanes stride over the neuron's synapses,
 * a fixed butterfly sums them (so results do not depend on scheduling),
 * lane 0 does the neuron arithmetic.
 */
#include <hip/hip_runtime.h>

#include "float_mode.h"

namespace {

constexpr unsigned BLOCK = 256;
constexpr unsigned WAVES = BLOCK / 64;
constexpr uint32_t IDX_MASK = 0x3ffffffu;
constexpr double H = 0.25;

__device__ inline double wave_sum(double x)
{
#pragma unroll
	for(int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
	return x;
}

__device__ inline double surrogate(double v)
{
	const double z = 1.0 + 10.0 * fabs(v - 30.0);
	return 1.0 / (z * z);
}

__global__ void __launch_bounds__(BLOCK)
forward_step_kernel(nemo_float_step_args a)
{
	const unsigned j = blockIdx.x * WAVES + threadIdx.x / 64;
	const unsigned lane = threadIdx.x & 63;
	if(j >= a.n) return;
	const unsigned t = a.t;
	const size_t n = a.n;
	double acc = 0.0;
	for(uint32_t k = a.rcm_ptr[j] + lane; k < a.rcm_ptr[j + 1]; k += 64) {
		const uint32_t sd = a.rcm_sd[k];
		const unsigned d = (sd >> 26) + 1;
		if(t >= d && a.S[size_t(t - d) * n + (sd & IDX_MASK)] != 0.0f) acc += double(a.wr[k]);
	}
	acc = wave_sum(acc);
	if(lane != 0) return;

	const double I = acc + double(a.x[size_t(t) * n + j]);
	const double v0 = a.v[j];
	const double u0 = a.u[j];
	if(a.record) {
		const size_t r = size_t(t - a.t0) * n + j;
		a.seg_v[r] = v0;
		a.seg_u[r] = u0;
		a.seg_I[r] = I;
	}
	if(a.ckpt && t % a.K == 0) {
		a.ckpt[size_t(t / a.K) * 2 * n + j] = v0;
		a.ckpt[size_t(t / a.K) * 2 * n + n + j] = u0;
	}
	const double A = a.pa[j], B = a.pb[j];
	double v = v0, u = u0;
	bool m = true;
#pragma unroll
	for(int k = 0; k < 4; ++k) {
		if(m) {
			v = v + H * (0.04 * v * v + 5.0 * v + 140.0 - u + I);
			u = u + H * A * (B * v - u);
			m = v < 30.0;
		}
	}
	const bool s = v >= 30.0;
	if(s) {
		v = double(a.pc[j]);
		u = u + double(a.pd[j]);
	}
	a.v[j] = v;
	a.u[j] = u;
	if(a.store) {
		a.S[size_t(t) * n + j] = s ? 1.0f : 0.0f;
		a.V[size_t(t) * n + j] = float(v);
	}
}

__global__ void __launch_bounds__(BLOCK)
backward_step_kernel(nemo_float_back_args a)
{
	const unsigned i = blockIdx.x * WAVES + threadIdx.x / 64;
	const unsigned lane = threadIdx.x & 63;
	if(i >= a.n) return;
	const unsigned tau = a.t;
	const size_t n = a.n;
	const unsigned R = a.ring;
	const bool fired = a.S[size_t(tau) * n + i] != 0.0f;

	/* adjoint of this step's spike from the steps it reaches (transposed
	 * delivery), and the weight gradient of the synapses it drove */
	double acc = 0.0;
	for(uint32_t k = a.fcm_ptr[i] + lane; k < a.fcm_ptr[i + 1]; k += 64) {
		const uint32_t td = a.fcm_td[k];
		const unsigned d = (td >> 26) + 1;
		if(tau + d < a.T) {
			const double l = a.LI[size_t((tau + d) % R) * n + (td & IDX_MASK)];
			acc += double(a.wf[k]) * l;
			if(fired) a.gw[k] = float(double(a.gw[k]) + l);
		}
	}
	acc = wave_sum(acc);
	if(lane != 0) return;

	const size_t r = size_t(tau) * n + i;
	const double lS = (a.gS ? double(a.gS[r]) : 0.0) + acc;
	double lv = a.LV[i] + (a.gV ? double(a.gV[r]) : 0.0); /* d L / d v(tau+1) */
	double lu = a.LU[i];                                  /* d L / d u(tau+1) */

	/* recompute the step's substeps from the recorded state and input */
	const size_t q = size_t(tau - a.t0) * n + i;
	const double I = a.seg_I[q];
	const double A = a.pa[i], B = a.pb[i];
	double vk[5], uk[5];
	bool mk[4];
	vk[0] = a.seg_v[q];
	uk[0] = a.seg_u[q];
	bool m = true;
#pragma unroll
	for(int k = 0; k < 4; ++k) {
		mk[k] = m;
		if(m) {
			vk[k + 1] = vk[k] + H * (0.04 * vk[k] * vk[k] + 5.0 * vk[k] + 140.0 - uk[k] + I);
			uk[k + 1] = uk[k] + H * A * (B * vk[k + 1] - uk[k]);
			m = vk[k + 1] < 30.0;
		} else {
			vk[k + 1] = vk[k];
			uk[k + 1] = uk[k];
		}
	}
	const double v4 = vk[4];
	const double s = v4 >= 30.0 ? 1.0 : 0.0;

	/* reset v' = v4 + s (c - v4), u' = u4 + s d with s constant; the spike
	 * s = H(v4 - 30) reaches v4 through the surrogate */
	double gd = lu * s;
	double Lv = lv * (1.0 - s) + lS * surrogate(v4);
	double Lu = lu;

	/* substeps in reverse; a masked substep is the identity */
	double lI = 0.0, ga = 0.0, gb = 0.0;
#pragma unroll
	for(int k = 3; k >= 0; --k) {
		if(!mk[k]) continue;
		const double Lv_tot = Lv + Lu * H * A * B;
		ga += Lu * H * (B * vk[k + 1] - uk[k]);
		gb += Lu * H * A * vk[k + 1];
		const double nLu = Lu * (1.0 - H * A) - Lv_tot * H;
		Lv = Lv_tot * (1.0 + H * (0.08 * vk[k] + 5.0));
		Lu = nLu;
		lI += Lv_tot * H;
	}
	a.LV[i] = Lv;
	a.LU[i] = Lu;
	a.LI[size_t(tau % R) * n + i] = lI;
	if(a.gx) a.gx[r] = float(lI);
	a.ga[i] += float(ga);
	a.gb[i] += float(gb);
	a.gd[i] += float(gd);
}

__global__ void gather_kernel(float* out, const float* in, const uint32_t* idx, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) out[k] = in[idx[k]];
}

__global__ void scatter_kernel(float* out, const float* in, const uint32_t* idx, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) out[idx[k]] = in[k];
}

__global__ void widen_kernel(double* out, const float* in, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) out[k] = double(in[k]);
}

__global__ void narrow_kernel(float* out, const double* in, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) out[k] = float(in[k]);
}

__global__ void fill_kernel(float* p, float v, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) p[k] = v;
}

__global__ void fill_double_kernel(double* p, double v, size_t n)
{
	const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
	if(k < n) p[k] = v;
}

inline unsigned blocks_for(size_t n)
{
	return unsigned((n + BLOCK - 1) / BLOCK);
}

} // namespace

extern "C" {

hipError_t nemo_launch_float_forward_step(hipStream_t st, const nemo_float_step_args* a)
{
	if(a->n == 0) return hipSuccess;
	if(a->record && (!a->seg_v || !a->seg_u || !a->seg_I || a->t < a->t0)) return hipErrorInvalidValue;
	if(a->ckpt && a->K == 0) return hipErrorInvalidValue;
	hipLaunchKernelGGL(forward_step_kernel, dim3(unsigned((a->n + WAVES - 1) / WAVES)), dim3(BLOCK), 0, st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_float_backward_step(hipStream_t st, const nemo_float_back_args* a)
{
	if(a->n == 0) return hipSuccess;
	if(a->ring == 0 || a->t >= a->T || a->t < a->t0) return hipErrorInvalidValue;
	hipLaunchKernelGGL(backward_step_kernel, dim3(unsigned((a->n + WAVES - 1) / WAVES)), dim3(BLOCK), 0, st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_gather(hipStream_t st, float* out, const float* in, const uint32_t* idx, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, out, in, idx, n);
	return hipGetLastError();
}

hipError_t nemo_launch_scatter(hipStream_t st, float* out, const float* in, const uint32_t* idx, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(scatter_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, out, in, idx, n);
	return hipGetLastError();
}

hipError_t nemo_launch_fill_double(hipStream_t st, double* p, double v, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(fill_double_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, p, v, n);
	return hipGetLastError();
}

hipError_t nemo_launch_narrow(hipStream_t st, float* out, const double* in, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(narrow_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, out, in, n);
	return hipGetLastError();
}

hipError_t nemo_launch_widen(hipStream_t st, double* out, const float* in, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(widen_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, out, in, n);
	return hipGetLastError();
}

hipError_t nemo_launch_fill(hipStream_t st, float* p, float v, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, st, p, v, n);
	return hipGetLastError();
}

} // extern "C"
