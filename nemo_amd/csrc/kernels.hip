/*
 * The backend's own gfx950 kernels: the per-step propagate kernel (binned
 * spike delivery), the STDP-only kernel, applyStdp and stimulus helpers.
 * Neuron updates, which also gather the binned input and usually run the
 * STDP accumulation, live in the model plugins (plugins/<Model>.hip).
 *
 * Data layout in HBM.  Device indices number neurons type by type, in
 * insertion order within a type (NeMo's simulation order,
 * src/nemo/cpu/Simulation.cpp:60-77), each type starting at a multiple of
 * NEMO_TILE (256):
 *   fcm        int2[S]     synapse rows (source, delay) back to back; row
 *                          (n, d) = [rowptr[n*D+d-1], rowptr[n*D+d]); entries
 *                          {target, Q11.20 weight}, 8 bytes, so consecutive
 *                          lanes read consecutive 8-byte words of a row
 *   fired_bits uint64[128][N/64]  one firing bitmap per step (ring)
 *   hist       uint64[2][N] firing history (bit k: fired k steps ago); plane
 *                          s % 2 holds the history after step s
 *   stdp_snap  uint64[N][2] two history snapshots per neuron (STDP batches)
 *   bins       int2[2][B][CAP] each delivery block's events, sorted by tile,
 *                          for steps of even and odd parity
 *   bin_table  uint32[2][T][B] count << 16 | offset of tile t in block b's
 *                          bins (entries of the blocks whose window holds t)
 *   acc        int64[2][2][N] overflow accumulators (events beyond CAP)
 *   rcm_*      per target, its plastic incoming synapses (STDP only)
 */
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>

#include "kernels.h"
#include "glibc_math.h"
#include "plugin_device.h"
#include "propagate_device.h"

namespace {

/* threads of a propagate (delivery or STDP) block */
constexpr unsigned PBLOCK = NEMO_PROPAGATE_BLOCK;

using nemo_dev::stdp_wrap_add;

/* delivery blocks, then (when fused into this launch) STDP blocks; at least
 * 6 waves per SIMD (80 VGPRs), so that three 512-thread blocks share a CU */
__global__ void __launch_bounds__(PBLOCK, 6)
propagate_kernel(nemo_propagate_args a)
{
	extern __shared__ uint64_t lds_dyn[];
	unsigned char* lds = reinterpret_cast<unsigned char*>(lds_dyn);
	const uint64_t t = a.clock->base_step + a.step_offset;
	if(blockIdx.x < a.deliver_blocks) {
		nemo_dev::deliver_block<PBLOCK>(a, t, blockIdx.x, lds);
	} else {
		nemo_dev::stdp_step<PBLOCK>(a, t, blockIdx.x - a.deliver_blocks, lds);
	}
}

/* Both delivery sets of one step in one launch: the first set's blocks, the
 * second's, then (when fused into this launch) STDP blocks of the first
 * set's arguments */
__global__ void __launch_bounds__(PBLOCK, 6)
propagate_sets_kernel(nemo_propagate_args a, nemo_propagate_args a2)
{
	extern __shared__ uint64_t lds_dyn[];
	unsigned char* lds = reinterpret_cast<unsigned char*>(lds_dyn);
	const uint64_t t = a.clock->base_step + a.step_offset;
	if(blockIdx.x < a.deliver_blocks) {
		nemo_dev::deliver_block<PBLOCK>(a, t, blockIdx.x, lds);
	} else if(blockIdx.x < a.deliver_blocks + a2.deliver_blocks) {
		nemo_dev::deliver_block<PBLOCK>(a2, t, blockIdx.x - a.deliver_blocks, lds);
	} else {
		nemo_dev::stdp_step<PBLOCK>(a, t, blockIdx.x - a.deliver_blocks - a2.deliver_blocks, lds);
	}
}

/* STDP blocks only */
__global__ void __launch_bounds__(PBLOCK)
stdp_kernel(nemo_propagate_args a)
{
	extern __shared__ uint64_t lds_dyn[];
	nemo_dev::stdp_step<PBLOCK>(a, a.clock->base_step + a.step_offset, blockIdx.x,
			reinterpret_cast<unsigned char*>(lds_dyn));
}

/* The pending STDP accumulation, one block per group of stdp_words bitmap
 * words (stdp_batch_block) */
__global__ void __launch_bounds__(NEMO_STDP_BATCH_BLOCK, NEMO_STDP_BATCH_WAVES)
stdp_batch_kernel(nemo_propagate_args a)
{
	extern __shared__ uint64_t lds_dyn[];
	nemo_dev::stdp_batch_block<NEMO_STDP_BATCH_BLOCK>(a, blockIdx.x, reinterpret_cast<unsigned char*>(lds_dyn));
}

/* The same as a flush, which nothing runs beside: wider blocks with more
 * loads in flight per thread */
__global__ void __launch_bounds__(PBLOCK)
stdp_flush_kernel(nemo_propagate_args a)
{
	extern __shared__ uint64_t lds_dyn[];
	nemo_dev::stdp_batch_block<PBLOCK, nemo_dev::STDP_FLUSH_ILP>(a, blockIdx.x,
			reinterpret_cast<unsigned char*>(lds_dyn));
}

/* A history snapshot after step base + offset into the free slot (the one
 * the last snapshot is not in) */
__global__ void __launch_bounds__(NEMO_BLOCK)
stdp_snapshot_kernel(nemo_propagate_args a, uint64_t offset)
{
	const uint64_t s = a.clock->base_step + offset;
	const unsigned slot = unsigned(a.clock->stdp_snaps + 1) & 1;
	const size_t n = a.fired_words * 64;
	for(size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x; i < n; i += size_t(gridDim.x) * NEMO_BLOCK) {
		a.stdp_snap[2 * i + slot] = a.hist[NEMO_HIST_INDEX(i, s, a.hist_pitch)];
	}
}

/* ... after which the steps since the last snapshot are pending; and the
 * step base moves on by advance (the last node of a graph of steps, so that
 * the next graph needs no copy of its base from the host) */
__global__ void stdp_snapshot_done_kernel(nemo_hip_clock* clock, uint64_t offset, uint64_t advance)
{
	const uint64_t base = clock->base_step;
	const uint64_t s = base + offset;
	clock->stdp_first = clock->stdp_last + 1;
	clock->stdp_count = s - clock->stdp_last;
	clock->stdp_last = s;
	clock->stdp_snaps += 1;
	if(advance) clock->base_step = base + advance;
}

/* ... and after a flush of the pending batch: nothing is pending, and the
 * accumulation of step base + offset has run */
__global__ void stdp_flush_done_kernel(nemo_hip_clock* clock, uint64_t offset)
{
	clock->stdp_count = 0;
	clock->stdp_flushed = clock->base_step + offset;
}

/* applyStdp in forward order: every synapse of the FCM that is plastic
 * (rcm_of[j] its accumulator) takes its weight change; reading the weights in
 * order and the accumulators scattered costs one random read per synapse,
 * where the reverse order costs a random read and write of the weight.  The
 * accumulators are cleared afterwards by a separate fill. */
__global__ void __launch_bounds__(NEMO_BLOCK)
apply_stdp_kernel(nemo_apply_args a)
{
	/* APPLY_ILP synapses per thread at a time: their index and weight loads,
	 * then their accumulator gathers, are in flight together */
	constexpr unsigned APPLY_ILP = 4;
	const size_t stride = size_t(gridDim.x) * NEMO_BLOCK * APPLY_ILP;
	for(size_t j0 = size_t(blockIdx.x) * NEMO_BLOCK * APPLY_ILP + threadIdx.x; j0 < a.fcm_count; j0 += stride) {
		uint32_t k[APPLY_ILP];
		int32_t w_old[APPLY_ILP], acc[APPLY_ILP];
#pragma unroll
		for(unsigned u = 0; u < APPLY_ILP; ++u) {
			const size_t j = j0 + u * NEMO_BLOCK;
			k[u] = j < a.fcm_count ? a.rcm_of[j] : ~0u;
			w_old[u] = j < a.fcm_count ? a.fcm[j].y : 0;
		}
#pragma unroll
		for(unsigned u = 0; u < APPLY_ILP; ++u) acc[u] = k[u] != ~0u && a.reward != 0 ? a.rcm_acc[k[u]] : 0;
#pragma unroll
		for(unsigned u = 0; u < APPLY_ILP; ++u) {
			if(k[u] == ~0u) continue;
			const int32_t dw = int32_t((int64_t(a.reward) * int64_t(acc[u])) >> a.fbits);
			int32_t w_new = 0;
			if(w_old[u] > 0) {
				const int32_t x = stdp_wrap_add(w_old[u], dw);
				w_new = min(a.maxE, max(x, a.minE));
			} else if(w_old[u] < 0) {
				const int32_t x = stdp_wrap_add(w_old[u], dw);
				w_new = min(a.minI, max(x, a.maxI));
			}
			if(w_new != w_old[u]) a.fcm[j0 + u * NEMO_BLOCK].y = w_new;
		}
	}
}

/* rcm_of[rcm_fwd[k]] = k for every plastic synapse k (rcm_of filled with ~0 before) */
__global__ void __launch_bounds__(NEMO_BLOCK)
invert_rcm_kernel(const uint32_t* rcm_fwd, uint32_t* rcm_of, size_t count)
{
	for(size_t k = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x; k < count; k += size_t(gridDim.x) * NEMO_BLOCK) {
		rcm_of[rcm_fwd[k]] = uint32_t(k);
	}
}

/* ------------------------------------------------------ firing readout */

/* Neurons fired in step t0 + blockIdx.x (its firing-bitmap ring slot) */
__global__ void __launch_bounds__(NEMO_BLOCK)
fired_count_kernel(const uint64_t* bits, size_t words, uint64_t t0, uint32_t* counts)
{
	const uint64_t* b = bits + ((t0 + blockIdx.x) & (NEMO_STEP_RING - 1)) * words;
	uint32_t c = 0;
	for(size_t w = threadIdx.x; w < words; w += NEMO_BLOCK) c += uint32_t(__popcll(b[w]));
	uint32_t total;
	nemo_dev::block_exclusive_scan<NEMO_BLOCK>(c, &total);
	if(threadIdx.x == 0) counts[blockIdx.x] = total;
}

/* The neurons fired in step t0 + blockIdx.x, in ascending device index, as
 * l2g's indices, written after those of the steps before it (counts) */
__global__ void __launch_bounds__(NEMO_BLOCK)
fired_list_kernel(const uint64_t* bits, size_t words, uint64_t t0, const uint32_t* counts, const uint32_t* l2g,
		uint32_t* list)
{
	__shared__ uint32_t s_off;
	if(threadIdx.x < 64) {
		uint32_t v = 0;
		for(unsigned s = threadIdx.x; s < blockIdx.x; s += 64) v += counts[s];
#pragma unroll
		for(int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
		if(threadIdx.x == 0) s_off = v;
	}
	__syncthreads();
	const uint64_t* b = bits + ((t0 + blockIdx.x) & (NEMO_STEP_RING - 1)) * words;
	/* a contiguous run of words per thread keeps the list in device order */
	const size_t per = (words + NEMO_BLOCK - 1) / NEMO_BLOCK;
	const size_t w0 = min(words, size_t(threadIdx.x) * per), w1 = min(words, w0 + per);
	uint32_t c = 0;
	for(size_t w = w0; w < w1; ++w) c += uint32_t(__popcll(b[w]));
	uint32_t total;
	uint32_t pos = s_off + nemo_dev::block_exclusive_scan<NEMO_BLOCK>(c, &total);
	for(size_t w = w0; w < w1; ++w) {
		uint64_t x = b[w];
		while(x) {
			const unsigned bit = unsigned(__builtin_ctzll(x));
			x &= x - 1;
			list[pos++] = l2g[w * 64 + bit];
		}
	}
}

/* ------------------------------------------- externally given firing */

/* The firing of step s given from outside (Brian's propagate,
 * cpu/Simulation.cpp:115-137): neuron ng fired if flags[sim] is nonzero for
 * its simulation index sim = dev_to_sim[ng] (~0 for padding); its history
 * and the step's bitmap word are written as an update would. */
__global__ void __launch_bounds__(NEMO_TILE)
external_fire_kernel(nemo_hip_step_buffers b, const uint32_t* flags, const uint32_t* dev_to_sim)
{
	const unsigned ng = blockIdx.x * NEMO_TILE + threadIdx.x;
	const uint32_t sim = dev_to_sim[ng];
	const bool valid = sim != ~0u;
	const uint64_t h = nemo_dev::load_history(b, ng, valid);
	nemo_dev::record_firing(b, ng, valid, valid && flags[sim] != 0u, h);
}

/* The synaptic input of a step, gathered per tile as the neuron updates
 * gather it (overflow accumulators cleared), converted to float and written
 * in simulation order: E[sim], I[sim]. */
__global__ void __launch_bounds__(NEMO_TILE)
gather_currents_kernel(nemo_hip_step_buffers b, const uint32_t* dev_to_sim, float* E, float* I)
{
	nemo_dev::UpdateLds& L = nemo_dev::update_lds<false>();
	nemo_hip_neuron_group g;
	g.start = 0;
	g.count = gridDim.x * NEMO_TILE;
	g.pitch = 0;
	g.param = nullptr;
	g.state = nullptr;
	g.rng = nullptr;
	g.scratch = nullptr;
	nemo_dev::gather_input(b, g, L);
	const unsigned ng = blockIdx.x * NEMO_TILE + threadIdx.x;
	float e, i;
	nemo_dev::take_current(b, L.input, ng, e, i);
	const uint32_t sim = dev_to_sim[ng];
	if(sim != ~0u) {
		E[sim] = e;
		I[sim] = i;
	}
}

/* ------------------------------------------------------------ stimulus */

__global__ void fstim_kernel(uint32_t* bits, const uint32_t* idx, size_t n, int set)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) {
		const uint32_t v = idx[i];
		if(set) atomicOr(&bits[v >> 5], 1u << (v & 31));
		else atomicAnd(&bits[v >> 5], ~(1u << (v & 31)));
	}
}

__global__ void istim_kernel(float* ext, const uint32_t* idx, const float* val, size_t n)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) ext[idx[i]] = val[i];
}

__global__ void eval_math_kernel(int fn, const float* in, float* out, size_t n)
{
	const size_t i = size_t(blockIdx.x) * NEMO_BLOCK + threadIdx.x;
	if(i < n) {
		const float x = in[i];
		out[i] = fn == 0 ? nemo_math::logf_exact(x) : fn == 1 ? nemo_math::sinf_exact(x) : nemo_math::expf_exact(x);
	}
}

} // namespace

extern "C" {

size_t nemo_propagate_lds_bytes(unsigned max_window, unsigned bin_capacity)
{
	return nemo_dev::PropagateLds<PBLOCK>::delivery_bytes(bin_capacity, max_window);
}

static bool stdp_batch_args_valid(const nemo_propagate_args* a, unsigned stdp_blocks)
{
	return a->stdp_snap && a->stdp.post_window <= 63 && a->stdp_words > 0 && a->stdp_words <= NEMO_STDP_BATCH_MAX_WORDS
		&& size_t(stdp_blocks) * a->stdp_words >= a->fired_words && a->stdp.rcm_ptr;
}

hipError_t nemo_launch_stdp_batch(hipStream_t st, const nemo_propagate_args* a, unsigned stdp_blocks)
{
	if(stdp_blocks == 0) return hipSuccess;
	if(!stdp_batch_args_valid(a, stdp_blocks)) return hipErrorInvalidValue;
	hipLaunchKernelGGL(stdp_batch_kernel, dim3(stdp_blocks), dim3(NEMO_STDP_BATCH_BLOCK),
			nemo_dev::PropagateLds<NEMO_STDP_BATCH_BLOCK>::stdp_bytes(), st, *a);
	return hipGetLastError();
}

hipError_t nemo_launch_stdp_flush(hipStream_t st, const nemo_propagate_args* a, unsigned stdp_blocks, uint64_t offset)
{
	if(stdp_blocks == 0) return hipSuccess;
	if(!stdp_batch_args_valid(a, stdp_blocks)) return hipErrorInvalidValue;
	hipLaunchKernelGGL(stdp_flush_kernel, dim3(stdp_blocks), dim3(PBLOCK), nemo_dev::PropagateLds<PBLOCK>::stdp_bytes(),
			st, *a);
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return e;
	hipLaunchKernelGGL(stdp_flush_done_kernel, dim3(1), dim3(1), 0, st, const_cast<nemo_hip_clock*>(a->clock), offset);
	return hipGetLastError();
}

hipError_t nemo_launch_stdp_snapshot(hipStream_t st, const nemo_propagate_args* a, uint64_t offset, uint64_t advance)
{
	if(!a->stdp_snap || !a->hist) return hipErrorInvalidValue;
	const size_t n = a->fired_words * 64;
	const unsigned blocks = unsigned(std::min<size_t>(1024, (n + NEMO_BLOCK - 1) / NEMO_BLOCK));
	hipLaunchKernelGGL(stdp_snapshot_kernel, dim3(blocks), dim3(NEMO_BLOCK), 0, st, *a, offset);
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return e;
	hipLaunchKernelGGL(stdp_snapshot_done_kernel, dim3(1), dim3(1), 0, st, const_cast<nemo_hip_clock*>(a->clock),
			offset, advance);
	return hipGetLastError();
}

size_t nemo_colaunch_lds_bytes(unsigned max_window, unsigned bin_capacity, int delivery)
{
	/* the update blocks of the launch take their LDS from the same buffer */
	return std::max(nemo_dev::PropagateLds<NEMO_COLAUNCH_BLOCK, nemo_dev::CARRIED_EVENT_ILP>::bytes(delivery != 0,
					bin_capacity, max_window),
			sizeof(nemo_dev::UpdateLds));
}

size_t nemo_propagate_static_lds(void)
{
	hipFuncAttributes fa;
	if(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(propagate_kernel)) != hipSuccess) return 32768;
	return fa.sharedSizeBytes;
}

hipError_t nemo_launch_propagate(hipStream_t st, const nemo_propagate_args* a, unsigned stdp_blocks, int fuse_stdp,
		hipEvent_t start, hipEvent_t stop)
{
	if(a->deliver_blocks > NEMO_MAX_BIN_BLOCKS) return hipErrorInvalidValue;
	if(a->deliver_blocks && (a->D == 0 || a->D > NEMO_MAX_DELAY || a->bin_capacity == 0
				|| a->bin_capacity > 32768 || a->max_window == 0 || a->max_window > a->tiles
				|| !a->pair_bounds || !a->windows || a->d_first == 0 || a->d_count == 0
				|| a->d_first + a->d_count - 1 > a->D || a->block_base + a->deliver_blocks > a->table_columns
				|| a->table_columns > NEMO_MAX_BIN_BLOCKS)) {
		return hipErrorInvalidValue;
	}
	if(stdp_blocks && (a->stdp.post_window > 63 || a->stdp_words == 0 || a->stdp_words > NEMO_STDP_MAX_WORDS
				|| size_t(stdp_blocks) * a->stdp_words < a->fired_words || !a->stdp.rcm_ptr || !a->hist)) {
		return hipErrorInvalidValue;
	}
	const size_t lds = nemo_dev::PropagateLds<PBLOCK>::bytes(a->deliver_blocks > 0, a->bin_capacity, a->max_window);
	if(fuse_stdp) {
		const unsigned blocks = a->deliver_blocks + stdp_blocks;
		if(blocks == 0) return hipSuccess;
		hipExtLaunchKernelGGL(propagate_kernel, dim3(blocks), dim3(PBLOCK), lds, st, start, stop, 0, *a);
		return hipGetLastError();
	}
	if(stdp_blocks) {
		hipExtLaunchKernelGGL(stdp_kernel, dim3(stdp_blocks), dim3(PBLOCK), nemo_dev::PropagateLds<PBLOCK>::stdp_bytes(),
				st, start, a->deliver_blocks ? nullptr : stop, 0, *a);
		const hipError_t e = hipGetLastError();
		if(e != hipSuccess) return e;
	}
	if(a->deliver_blocks) {
		hipExtLaunchKernelGGL(propagate_kernel, dim3(a->deliver_blocks), dim3(PBLOCK), lds, st,
				stdp_blocks ? nullptr : start, stop, 0, *a);
		return hipGetLastError();
	}
	return hipSuccess;
}

hipError_t nemo_launch_propagate_sets(hipStream_t st, const nemo_propagate_args* a, const nemo_propagate_args* a2,
		unsigned stdp_blocks, hipEvent_t start, hipEvent_t stop)
{
	if(!a2->deliver_blocks) return nemo_launch_propagate(st, a, stdp_blocks, 1, start, stop);
	if(!a->deliver_blocks) {
		if(stdp_blocks) return hipErrorInvalidValue; /* the STDP blocks take the first set's arguments */
		return nemo_launch_propagate(st, a2, 0, 1, start, stop);
	}
	if(a->step_offset != a2->step_offset || a->clock != a2->clock || a->bin_capacity != a2->bin_capacity
			|| a->max_window != a2->max_window || a->table_columns != a2->table_columns
			|| a->deliver_blocks + a2->deliver_blocks > NEMO_MAX_BIN_BLOCKS) {
		return hipErrorInvalidValue;
	}
	for(const nemo_propagate_args* x : {a, a2}) {
		if(x->D == 0 || x->D > NEMO_MAX_DELAY || x->bin_capacity == 0 || x->bin_capacity > 32768
				|| x->max_window == 0 || x->max_window > x->tiles || !x->pair_bounds || !x->windows
				|| x->d_first == 0 || x->d_count == 0 || x->d_first + x->d_count - 1 > x->D
				|| x->block_base + x->deliver_blocks > x->table_columns || x->table_columns > NEMO_MAX_BIN_BLOCKS) {
			return hipErrorInvalidValue;
		}
	}
	if(stdp_blocks && (a->stdp.post_window > 63 || a->stdp_words == 0 || a->stdp_words > NEMO_STDP_MAX_WORDS
				|| size_t(stdp_blocks) * a->stdp_words < a->fired_words || !a->stdp.rcm_ptr || !a->hist)) {
		return hipErrorInvalidValue;
	}
	const size_t lds = nemo_dev::PropagateLds<PBLOCK>::bytes(true, a->bin_capacity, a->max_window);
	const unsigned blocks = a->deliver_blocks + a2->deliver_blocks + stdp_blocks;
	hipExtLaunchKernelGGL(propagate_sets_kernel, dim3(blocks), dim3(PBLOCK), lds, st, start, stop, 0, *a, *a2);
	return hipGetLastError();
}

hipError_t nemo_launch_apply_stdp(hipStream_t st, const nemo_apply_args* a)
{
	if(a->count == 0) return hipSuccess;
	size_t blocks = (a->fcm_count + 4 * NEMO_BLOCK - 1) / (4 * NEMO_BLOCK);
	if(blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(apply_stdp_kernel, dim3(unsigned(blocks)), dim3(NEMO_BLOCK), 0, st, *a);
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return e;
	return hipMemsetAsync(a->rcm_acc, 0, a->count * sizeof(int32_t), st);
}

hipError_t nemo_launch_invert_rcm(hipStream_t st, const uint32_t* rcm_fwd, uint32_t* rcm_of, size_t count,
		size_t fcm_count)
{
	hipError_t e = hipMemsetAsync(rcm_of, 0xff, fcm_count * sizeof(uint32_t), st);
	if(e != hipSuccess || count == 0) return e;
	size_t blocks = (count + NEMO_BLOCK - 1) / NEMO_BLOCK;
	if(blocks > 16384) blocks = 16384;
	hipLaunchKernelGGL(invert_rcm_kernel, dim3(unsigned(blocks)), dim3(NEMO_BLOCK), 0, st, rcm_fwd, rcm_of, count);
	return hipGetLastError();
}

hipError_t nemo_launch_external_fire(hipStream_t st, const nemo_hip_step_buffers* b, const uint32_t* flags,
		const uint32_t* dev_to_sim, unsigned tiles)
{
	if(tiles == 0) return hipSuccess;
	hipLaunchKernelGGL(external_fire_kernel, dim3(tiles), dim3(NEMO_TILE), 0, st, *b, flags, dev_to_sim);
	return hipGetLastError();
}

hipError_t nemo_launch_gather_currents(hipStream_t st, const nemo_hip_step_buffers* b, const uint32_t* dev_to_sim,
		float* E, float* I, unsigned tiles)
{
	if(tiles == 0) return hipSuccess;
	hipLaunchKernelGGL(gather_currents_kernel, dim3(tiles), dim3(NEMO_TILE), 0, st, *b, dev_to_sim, E, I);
	return hipGetLastError();
}

hipError_t nemo_launch_fired_list(hipStream_t st, const uint64_t* bits, size_t words, uint64_t t0, unsigned k,
		const uint32_t* l2g, uint32_t* counts, uint32_t* list)
{
	if(k == 0) return hipSuccess;
	if(k > NEMO_STEP_RING || !bits || !l2g || !counts || !list) return hipErrorInvalidValue;
	hipLaunchKernelGGL(fired_count_kernel, dim3(k), dim3(NEMO_BLOCK), 0, st, bits, words, t0, counts);
	const hipError_t e = hipGetLastError();
	if(e != hipSuccess) return e;
	hipLaunchKernelGGL(fired_list_kernel, dim3(k), dim3(NEMO_BLOCK), 0, st, bits, words, t0, counts, l2g, list);
	return hipGetLastError();
}

hipError_t nemo_launch_fstim(hipStream_t st, uint32_t* bits, const uint32_t* idx, size_t n, int set)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(fstim_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			bits, idx, n, set);
	return hipGetLastError();
}

hipError_t nemo_launch_istim(hipStream_t st, float* ext, const uint32_t* idx, const float* val, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(istim_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			ext, idx, val, n);
	return hipGetLastError();
}

hipError_t nemo_launch_eval_math(hipStream_t st, int fn, const float* in, float* out, size_t n)
{
	if(n == 0) return hipSuccess;
	hipLaunchKernelGGL(eval_math_kernel, dim3(unsigned((n + NEMO_BLOCK - 1) / NEMO_BLOCK)), dim3(NEMO_BLOCK), 0, st,
			fn, in, out, n);
	return hipGetLastError();
}

} // extern "C"
