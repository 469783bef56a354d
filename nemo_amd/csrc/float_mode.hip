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
 * A group of lanes per neuron: they stride over the neuron's synapses, a
 * fixed butterfly sums them (so results do not depend on scheduling), the
 * group's first lane does the neuron arithmetic.  The forward step tests a
 * synapse's spike in a bitmap of recent steps ([T][pitch] words, bit j of
 * step t set when neuron j fired), which each block stages in LDS when the
 * last dmax steps' bitmaps are small; a block of 32 neurons writes its step's
 * bitmap word itself.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "float_mode.h"

namespace {

/* Backward: LANES lanes per neuron, BLOCK threads per block. */
constexpr unsigned BLOCK = 256;
#ifndef NEMO_FLOAT_LANES
#define NEMO_FLOAT_LANES 32
#endif
constexpr unsigned LANES = NEMO_FLOAT_LANES;
constexpr unsigned GROUPS = BLOCK / LANES;
static_assert(LANES >= 8 && LANES <= 64 && (LANES & (LANES - 1)) == 0, "LANES: a power of two, 8..64");
/* Forward: FLANES lanes per neuron and 32 neurons per block, so that a
 * block's spikes make one word of the spike bitmap. */
#ifndef NEMO_FLOAT_FWD_LANES
#define NEMO_FLOAT_FWD_LANES 32
#endif
constexpr unsigned FLANES = NEMO_FLOAT_FWD_LANES;
constexpr unsigned FGROUPS = 32;
constexpr unsigned FBLOCK = FGROUPS * FLANES;
static_assert(FLANES >= 8 && FLANES <= 32 && (FLANES & (FLANES - 1)) == 0, "FLANES: a power of two, 8..32");
constexpr uint32_t IDX_MASK = 0x3ffffffu;
constexpr double H = 0.25;
/* synapses in flight per lane */
template<unsigned L> constexpr unsigned unroll_for() { return L >= 16 ? 256 / L : 16; }

/* x of the partner lane under a DPP control (one 32-bit half at a time):
 * quad_perm [1,0,3,2] pairs lanes 4k+0/1 and 4k+2/3, [2,3,0,1] pairs 4k+0/2
 * and 4k+1/3, row_half_mirror pairs i and 7-i in each 8, row_mirror pairs i
 * and 15-i in each 16. */
template<int CTRL>
__device__ inline double dpp_partner(double x)
{
	const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
	return __hiloint2double(hi, lo);
}
constexpr int DPP_QUAD_1032 = 0xb1;
constexpr int DPP_QUAD_2301 = 0x4e;
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

/* Sum over groups of L consecutive lanes (L a power of two up to 64), the
 * same fixed pairing in every call: within quads, eights and rows of 16 by
 * DPP, across rows by shuffles.  The first lane of a group (the one that
 * reads it) always adds in the same order, so a sum is reproducible bit for
 * bit. */
template<unsigned L>
__device__ inline double group_sum(double x)
{
	static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L: a power of two up to 64");
	if(L > 1) x += dpp_partner<DPP_QUAD_1032>(x);
	if(L > 2) x += dpp_partner<DPP_QUAD_2301>(x);
	if(L > 4) x += dpp_partner<DPP_ROW_HALF_MIRROR>(x);
	if(L > 8) x += dpp_partner<DPP_ROW_MIRROR>(x);
	if(L > 16) x += __shfl_xor(x, 16);
	if(L > 32) x += __shfl_xor(x, 32);
	return x;
}

__device__ inline double surrogate(double v)
{
	const double z = 1.0 + 10.0 * fabs(v - 30.0);
	return 1.0 / (z * z);
}

/* One step of the neuron from (v, u) with input I: four substeps while
 * v < 30, then the reset; returns the spike.  The forward and the replay
 * share it, so that the replay reproduces the forward's state exactly. */
__device__ inline bool neuron_step(double& v, double& u, double I, double A, double B, float C, float D)
{
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
		v = double(C);
		u = u + double(D);
	}
	return s;
}

/* Spike test of synapse (source src, delay d) at step t: bit src of the
 * bitmap of step t - d, from the block's staged rows (row d-1 holds step
 * t-d) or from global memory. */
template<bool STAGED>
__device__ inline bool spiked(const nemo_float_step_args& a, const uint32_t* stage, unsigned t, unsigned d, uint32_t src)
{
	const uint32_t w = STAGED ? stage[(d - 1) * a.pitch + (src >> 5)] : a.bits[size_t(t - d) * a.pitch + (src >> 5)];
	return (w >> (src & 31)) & 1u;
}

template<bool STAGED>
__global__ void __launch_bounds__(FBLOCK)
forward_step_kernel(nemo_float_step_args a)
{
	constexpr unsigned UNROLL = unroll_for<FLANES>();
	extern __shared__ uint4 stage4[];
	__shared__ uint32_t fired;
	const unsigned t = a.t;
	const size_t n = a.n;
	const unsigned g = threadIdx.x / FLANES;
	const unsigned j = blockIdx.x * FGROUPS + g;
	const unsigned lane = threadIdx.x % FLANES;
	const bool live = j < n;
	uint32_t beg = 0, end = 0;
	float xj = 0.0f, C = 0.0f, Dd = 0.0f;
	double v0 = 0.0, u0 = 0.0, A = 0.0, B = 0.0;
	if(live) {
		/* the neuron's own operands, loaded before the staging and the
		 * synapses so that their latency overlaps */
		beg = a.rcm_ptr[j];
		end = a.rcm_ptr[j + 1];
		xj = a.x[size_t(t) * n + j];
		v0 = a.v[j];
		u0 = a.u[j];
		A = a.pa[j];
		B = a.pb[j];
		C = a.pc[j];
		Dd = a.pd[j];
	}
	/* UNROLL synapses per lane in flight; the sum runs in ascending synapse
	 * order per lane.  The first batch is issued before the staging so that
	 * both are in flight together. */
	uint32_t sd[UNROLL];
	float w[UNROLL];
	uint32_t k0 = beg + lane;
#pragma unroll
	for(unsigned q = 0; q < UNROLL; ++q) {
		const uint32_t k = k0 + FLANES * q;
		sd[q] = k < end ? a.rcm_sd[k] : 0u;
		w[q] = k < end ? a.wr[k] : 0.0f;
	}
	if(STAGED) {
		/* bitmaps of steps t-1 .. t-min(dmax, t), pitch a multiple of 4 words */
		const unsigned q4 = a.pitch / 4;
		const unsigned total = min(a.dmax, t) * q4;
		for(unsigned e = threadIdx.x; e < total; e += FBLOCK) {
			const unsigned r = e / q4;
			stage4[e] = reinterpret_cast<const uint4*>(a.bits + size_t(t - 1 - r) * a.pitch)[e - r * q4];
		}
	}
	if(threadIdx.x == 0) fired = 0;
	__syncthreads();
	const uint32_t* stage = reinterpret_cast<const uint32_t*>(stage4);

	bool s = false;
	if(live) {
		double acc = 0.0;
		while(k0 < end) {
#pragma unroll
			for(unsigned q = 0; q < UNROLL; ++q) {
				const unsigned d = (sd[q] >> 26) + 1;
				if(k0 + FLANES * q < end && t >= d && spiked<STAGED>(a, stage, t, d, sd[q] & IDX_MASK))
					acc += double(w[q]);
			}
			k0 += FLANES * UNROLL;
#pragma unroll
			for(unsigned q = 0; q < UNROLL; ++q) {
				const uint32_t k = k0 + FLANES * q;
				sd[q] = k < end ? a.rcm_sd[k] : 0u;
				w[q] = k < end ? a.wr[k] : 0.0f;
			}
		}
		acc = group_sum<FLANES>(acc);
		if(lane == 0) {
			const double I = acc + double(xj);
			if(a.ckpt && t % a.K == 0) {
				a.ckpt[size_t(t / a.K) * 2 * n + j] = v0;
				a.ckpt[size_t(t / a.K) * 2 * n + n + j] = u0;
			}
			double v = v0, u = u0;
			s = neuron_step(v, u, I, A, B, C, Dd);
			a.v[j] = v;
			a.u[j] = u;
			if(a.store) {
				a.S[size_t(t) * n + j] = s ? 1.0f : 0.0f;
				a.V[size_t(t) * n + j] = float(v);
			}
		}
	}
	if(a.store) {
		/* the block's 32 neurons are word blockIdx.x of step t's bitmap */
		if(s) atomicOr(&fired, 1u << g);
		__syncthreads();
		if(threadIdx.x == 0) a.bits[size_t(t) * a.pitch + blockIdx.x] = fired;
	}
}

__global__ void __launch_bounds__(BLOCK)
backward_step_kernel(nemo_float_back_args a)
{
	const unsigned i = blockIdx.x * GROUPS + threadIdx.x / LANES;
	const unsigned lane = threadIdx.x % LANES;
	if(i >= a.n) return;
	const unsigned tau = a.t;
	const size_t n = a.n;
	const unsigned R = a.ring;
	const uint32_t beg = a.fcm_ptr[i], end = a.fcm_ptr[i + 1];
	const size_t r = size_t(tau) * n + i;
	const bool fired = a.S[r] != 0.0f;
	/* the neuron's own operands, loaded before the synapses so that their
	 * latency overlaps the gathers' */
	const size_t sq = size_t(tau - a.t0) * n + i;
	const float gSr = a.gS ? a.gS[r] : 0.0f;
	const float gVr = a.gV ? a.gV[r] : 0.0f;
	const double LVi = a.LV[i], LUi = a.LU[i];
	const double I = a.seg_I[sq];
	const double sv = a.seg_v[sq], su = a.seg_u[sq];
	const double A = a.pa[i], B = a.pb[i];
	const float ga0 = a.ga[i], gb0 = a.gb[i], gd0 = a.gd[i];

	/* adjoint of this step's spike from the steps it reaches (transposed
	 * delivery), and the weight gradient of the synapses it drove; UNROLL
	 * synapses per lane in flight, summed in ascending synapse order */
	/* 4 in flight per lane keeps the kernel at 7 waves per SIMD; 8 would
	 * need more registers than 5 waves allow */
	constexpr unsigned UNROLL = 4;
	const unsigned tm = tau % R; /* slot of step tau + d: tm + d, less R once past it (d < R) */
	double acc = 0.0;
	for(uint32_t k0 = beg + lane; k0 < end; k0 += LANES * UNROLL) {
		uint32_t td[UNROLL];
		float w[UNROLL], g[UNROLL];
#pragma unroll
		for(unsigned q = 0; q < UNROLL; ++q) {
			const uint32_t k = k0 + LANES * q;
			td[q] = k < end ? a.fcm_td[k] : 0u;
			w[q] = k < end ? a.wf[k] : 0.0f;
			g[q] = k < end && fired ? a.gw[k] : 0.0f;
		}
		bool live[UNROLL];
		double l[UNROLL];
#pragma unroll
		for(unsigned q = 0; q < UNROLL; ++q) {
			const unsigned d = (td[q] >> 26) + 1;
			live[q] = k0 + LANES * q < end && tau + d < a.T;
			const unsigned slot = tm + d >= R ? tm + d - R : tm + d;
			l[q] = live[q] ? a.LI[size_t(slot) * n + (td[q] & IDX_MASK)] : 0.0;
		}
#pragma unroll
		for(unsigned q = 0; q < UNROLL; ++q) {
			if(!live[q]) continue;
			acc += double(w[q]) * l[q];
			if(fired) a.gw[k0 + LANES * q] = float(double(g[q]) + l[q]);
		}
	}
	acc = group_sum<LANES>(acc);
	if(lane != 0) return;

	const double lS = double(gSr) + acc;
	double lv = LVi + double(gVr); /* d L / d v(tau+1) */
	double lu = LUi;               /* d L / d u(tau+1) */

	/* recompute the step's substeps from the recorded state and input */
	double vk[5], uk[5];
	bool mk[4];
	vk[0] = sv;
	uk[0] = su;
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
	a.ga[i] = ga0 + float(ga);
	a.gb[i] = gb0 + float(gb);
	a.gd[i] = gd0 + float(gd);
}

/* Firing histories for the replay: hist[r][j] bit k set when neuron j
 * fired at step 64 (r - 1) + k; row 0 (steps -64..-1) and steps from T on
 * are zero. */
__global__ void __launch_bounds__(BLOCK)
pack_history_kernel(uint64_t* hist, unsigned rows, const float* S, unsigned n, unsigned T)
{
	const unsigned j = blockIdx.x * BLOCK + threadIdx.x;
	const unsigned r = blockIdx.y;
	if(j >= n || r >= rows) return;
	uint64_t word = 0;
	if(r > 0) {
		const unsigned t0 = 64 * (r - 1);
		for(unsigned k = 0; k < 64 && t0 + k < T; ++k)
			if(S[size_t(t0 + k) * n + j] != 0.0f) word |= uint64_t(1) << k;
	}
	hist[size_t(r) * n + j] = word;
}

/* Steps t0..t1-1 of the forward in one launch, from the state before t0 and
 * the recorded spikes: each step's entry state and input are recorded.  The
 * synaptic sums are those of forward_step_kernel, operation for operation
 * (same lanes per neuron, same synapse-to-lane assignment and order, same
 * butterfly), so the replayed state equals the forward's.  CH steps at a
 * time: each synapse's spikes over the CH steps come from at most two
 * history words. */
constexpr unsigned CH = 16;

__global__ void __launch_bounds__(BLOCK)
replay_kernel(nemo_float_replay_args a)
{
	constexpr unsigned UNROLL = 4;
	constexpr unsigned RGROUPS = BLOCK / FLANES;
	static_assert(CH <= FLANES, "lane c holds the input of step c of a chunk");
	const unsigned j = blockIdx.x * RGROUPS + threadIdx.x / FLANES;
	const unsigned lane = threadIdx.x % FLANES;
	if(j >= a.n) return;
	const size_t n = a.n;
	const uint32_t beg = a.rcm_ptr[j], end = a.rcm_ptr[j + 1];
	const double A = a.pa[j], B = a.pb[j];
	const float C = a.pc[j], Dd = a.pd[j];
	double v = a.v0[j], u = a.u0[j];
	for(unsigned c0 = a.t0; c0 < a.t1; c0 += CH) {
		const unsigned steps = min(CH, a.t1 - c0);
		const float xl = lane < steps ? a.x[size_t(c0 + lane) * n + j] : 0.0f;
		double acc[CH];
#pragma unroll
		for(unsigned q = 0; q < CH; ++q) acc[q] = 0.0;
		for(uint32_t k0 = beg + lane; k0 < end; k0 += FLANES * UNROLL) {
			uint32_t sd[UNROLL];
			float w[UNROLL];
#pragma unroll
			for(unsigned q = 0; q < UNROLL; ++q) {
				const uint32_t k = k0 + FLANES * q;
				sd[q] = k < end ? a.rcm_sd[k] : 0u;
				w[q] = k < end ? a.wr[k] : 0.0f;
			}
			/* the source's spikes at steps c0-d .. c0-d+CH-1: history bits
			 * from p = c0 - d + 64 (d <= 64, so p >= 0) */
			uint64_t lo[UNROLL], hi[UNROLL];
#pragma unroll
			for(unsigned q = 0; q < UNROLL; ++q) {
				const bool valid = k0 + FLANES * q < end;
				const unsigned p = c0 + 64 - ((sd[q] >> 26) + 1);
				const size_t row = size_t(p >> 6) * n + (sd[q] & IDX_MASK);
				lo[q] = valid ? a.hist[row] : 0;
				hi[q] = valid && (p & 63) + CH > 64 ? a.hist[row + n] : 0;
			}
#pragma unroll
			for(unsigned q = 0; q < UNROLL; ++q) {
				const unsigned o = (c0 + 64 - ((sd[q] >> 26) + 1)) & 63;
				const uint64_t win = o ? (lo[q] >> o) | (hi[q] << (64 - o)) : lo[q];
#pragma unroll
				for(unsigned c = 0; c < CH; ++c)
					if((win >> c) & 1) acc[c] += double(w[q]);
			}
		}
#pragma unroll
		for(unsigned c = 0; c < CH; ++c) {
			const double tot = group_sum<FLANES>(acc[c]);
			const float xc = __shfl(xl, c, FLANES);
			if(lane == 0 && c < steps) {
				const double I = tot + double(xc);
				const size_t rec = size_t(c0 + c - a.t0) * n + j;
				a.seg_v[rec] = v;
				a.seg_u[rec] = u;
				a.seg_I[rec] = I;
				neuron_step(v, u, I, A, B, C, Dd);
			}
		}
	}
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

/* The workspace of the allocation-free forward (see nemo_float_scratch).
 * Host code never passes its name to a kernel: the name is a host-side
 * stand-in, and the device address comes from hipGetSymbolAddress. */
static __device__ __align__(256) unsigned char scratch[NEMO_FLOAT_SCRATCH_BYTES];

extern "C" {

hipError_t nemo_float_scratch(void** ptr, size_t* bytes)
{
	static std::mutex lock;
	static void* addr[NEMO_FLOAT_SCRATCH_DEVICES] = {};
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return e;
	if(dev < 0 || dev >= NEMO_FLOAT_SCRATCH_DEVICES) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> held(lock);
	if(!addr[dev]) {
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess) return e;
		if(!p) return hipErrorInvalidSymbol;
		addr[dev] = p;
	}
	*ptr = addr[dev];
	*bytes = sizeof(scratch);
	return hipSuccess;
}

hipError_t nemo_launch_float_forward_step(hipStream_t st, const nemo_float_step_args* a)
{
	if(a->n == 0) return hipSuccess;
	if(a->ckpt && a->K == 0) return hipErrorInvalidValue;
	if(!a->bits || a->pitch % 4 || a->pitch < nemo_float_bitmap_pitch(a->n) || a->dmax == 0)
		return hipErrorInvalidValue;
	const dim3 grid(unsigned((a->n + FGROUPS - 1) / FGROUPS));
	if(a->stage) {
		const size_t lds = nemo_float_stage_bytes(a->n, a->dmax);
		if(lds > NEMO_FLOAT_STAGE_MAX) return hipErrorInvalidValue;
		hipLaunchKernelGGL(forward_step_kernel<true>, grid, dim3(FBLOCK), lds, st, *a);
	} else {
		hipLaunchKernelGGL(forward_step_kernel<false>, grid, dim3(FBLOCK), 0, st, *a);
	}
	return hipGetLastError();
}

unsigned nemo_float_bitmap_pitch(unsigned n)
{
	return ((n + 31) / 32 + 3) / 4 * 4;
}

size_t nemo_float_stage_bytes(unsigned n, unsigned dmax)
{
	return size_t(dmax) * nemo_float_bitmap_pitch(n) * sizeof(uint32_t);
}

hipError_t nemo_launch_pack_history(hipStream_t st, uint64_t* hist, unsigned rows, const float* S, unsigned n,
		unsigned T)
{
	if(n == 0 || rows == 0) return hipSuccess;
	if(rows < nemo_float_history_rows(T)) return hipErrorInvalidValue;
	hipLaunchKernelGGL(pack_history_kernel, dim3(blocks_for(n), rows), dim3(BLOCK), 0, st, hist, rows, S, n, T);
	return hipGetLastError();
}

unsigned nemo_float_history_rows(unsigned T)
{
	return (T + 63) / 64 + 2;
}

hipError_t nemo_launch_float_replay(hipStream_t st, const nemo_float_replay_args* a)
{
	if(a->n == 0 || a->t1 <= a->t0) return hipSuccess;
	if(!a->hist || a->hist_rows < nemo_float_history_rows(a->t1) || a->dmax == 0 || a->dmax > 64)
		return hipErrorInvalidValue;
	constexpr unsigned RGROUPS = BLOCK / FLANES;
	hipLaunchKernelGGL(replay_kernel, dim3(unsigned((a->n + RGROUPS - 1) / RGROUPS)), dim3(BLOCK), 0, st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_float_backward_step(hipStream_t st, const nemo_float_back_args* a)
{
	if(a->n == 0) return hipSuccess;
	if(a->ring == 0 || a->t >= a->T || a->t < a->t0) return hipErrorInvalidValue;
	hipLaunchKernelGGL(backward_step_kernel, dim3(unsigned((a->n + GROUPS - 1) / GROUPS)), dim3(BLOCK), 0, st, *a);
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
