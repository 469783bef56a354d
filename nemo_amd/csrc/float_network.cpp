/*
 * Host side of the float-weight mode: connectivity in forward (network)
 * order and reverse (by target) order, and the forward / checkpointed
 * adjoint step loops over the kernels of float_mode.hip.
 */
#include "backend.hpp"
#include "float_mode.h"

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

namespace nemo_hip {

FloatNetwork::FloatNetwork(const Network& net, int device) : m_device(device < 0 ? 0 : device)
{
	HIP_CHECK(hipSetDevice(m_device));
	unsigned types = 0;
	const Network::Group* grp = nullptr;
	for(unsigned t = 0; t < net.types.size(); ++t) {
		if(net.groups[t].gidx.empty()) continue;
		++types;
		grp = &net.groups[t];
		if(net.types[t].name != "Izhikevich") {
			throw Error(NEMO_API_UNSUPPORTED, "the float-weight mode supports Izhikevich neurons only");
		}
	}
	if(types != 1) {
		throw Error(NEMO_INVALID_INPUT, "the float-weight mode needs one neuron type");
	}
	m_n = grp->gidx.size();
	if(m_n >= (size_t(1) << 26)) throw Error(NEMO_API_UNSUPPORTED, "too many neurons for the float-weight mode");
	std::vector<char> seen(m_n, 0);
	for(unsigned g : grp->gidx) {
		if(g >= m_n || seen[g]) {
			throw Error(NEMO_INVALID_INPUT, "the float-weight mode needs neuron indices 0..N-1");
		}
		seen[g] = 1;
	}
	m_s = net.synapseCount;
	if(m_s >= (size_t(1) << 32)) throw Error(NEMO_API_UNSUPPORTED, "too many synapses for the float-weight mode");
	m_D = std::max(1u, net.maxDelay);

	/* forward order = network order: sources ascending, insertion order */
	std::vector<uint32_t> fptr(m_n + 1, 0), ftd(std::max<size_t>(1, m_s)), tgtOf(std::max<size_t>(1, m_s));
	std::vector<uint32_t> incoming(m_n + 1, 0);
	size_t k = 0;
	for(auto& kv : net.fcm) {
		const Network::Axon& ax = kv.second;
		if(kv.first >= m_n) {
			throw Error(NEMO_INVALID_INPUT, "Invalid synapse source neuron " + std::to_string(kv.first));
		}
		fptr[kv.first + 1] = uint32_t(ax.target.size());
		for(size_t i = 0; i < ax.target.size(); ++i, ++k) {
			const uint32_t tgt = ax.target[i];
			if(tgt >= m_n) throw Error(NEMO_INVALID_INPUT, "Invalid synapse target neuron " + std::to_string(tgt));
			ftd[k] = tgt | (uint32_t(ax.delay[i] - 1) << 26);
			tgtOf[k] = tgt;
			incoming[tgt + 1] += 1;
		}
	}
	for(size_t i = 0; i < m_n; ++i) fptr[i + 1] += fptr[i];
	/* reverse order: by target, then network order */
	for(size_t i = 0; i < m_n; ++i) incoming[i + 1] += incoming[i];
	std::vector<uint32_t> cursor(incoming.begin(), incoming.end() - 1);
	std::vector<uint32_t> rsd(std::max<size_t>(1, m_s)), ruser(std::max<size_t>(1, m_s));
	k = 0;
	for(auto& kv : net.fcm) {
		const Network::Axon& ax = kv.second;
		for(size_t i = 0; i < ax.target.size(); ++i, ++k) {
			const uint32_t pos = cursor[tgtOf[k]]++;
			rsd[pos] = kv.first | (uint32_t(ax.delay[i] - 1) << 26);
			ruser[pos] = uint32_t(k);
		}
	}
	auto upload = [](DeviceArray<uint32_t>& d, const std::vector<uint32_t>& h) {
		d.allocate(h.size());
		HIP_CHECK(hipMemcpy(d.ptr, h.data(), d.bytes(), hipMemcpyHostToDevice));
	};
	upload(m_fcmPtr, fptr);
	upload(m_fcmTd, ftd);
	upload(m_rcmPtr, incoming);
	upload(m_rcmSd, rsd);
	upload(m_rcmUser, ruser);
	m_wr.allocate(std::max<size_t>(1, m_s));
	m_v.allocate(m_n);
	m_u.allocate(m_n);
	m_LV.allocate(m_n);
	m_LU.allocate(m_n);
	m_LI.allocate(size_t(m_D + 1) * m_n);
	m_pitch = nemo_float_bitmap_pitch(unsigned(m_n));
	/* stage the recent spike bitmaps in LDS when they fit and cost fewer
	 * cache lines than a block's (32 neurons') spike gathers would */
	const size_t stage = nemo_float_stage_bytes(unsigned(m_n), m_D);
	m_stage = stage <= NEMO_FLOAT_STAGE_MAX && stage / 64 <= 32 * (m_s / std::max<size_t>(1, m_n));
	/* NEMO_HIP_FLOAT_STAGE=0 reads the bitmaps from global memory instead */
	if(const char* e = std::getenv("NEMO_HIP_FLOAT_STAGE")) {
		if(std::string(e) == "0") m_stage = false;
	}
}

void FloatNetwork::prepare(unsigned T, unsigned K)
{
	if(T == 0) throw Error(NEMO_INVALID_INPUT, "the float-weight mode needs at least one step");
	if(K == 0) throw Error(NEMO_INVALID_INPUT, "checkpoint interval must be positive");
	HIP_CHECK(hipSetDevice(m_device));
	const unsigned k = std::min(K, T);
	if(k > m_segK) {
		m_seg.allocate(3 * size_t(k) * m_n);
		m_segK = k;
	}
	if(T > m_bitsT) {
		m_bits.allocate(size_t(T) * m_pitch);
		m_bitsT = T;
	}
}

void FloatNetwork::forward(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a,
		const float* b, const float* c, const float* d, const float* v0, const float* u0, const float* x,
		float* S, float* V, double* ckpt)
{
	prepare(T, K);
	ForwardWorkspace ws;
	ws.wr = m_wr.ptr;
	ws.v = m_v.ptr;
	ws.u = m_u.ptr;
	ws.bits = m_bits.ptr;
	forwardWith(ws, st, T, K, w, a, b, c, d, v0, u0, x, S, V, ckpt);
}

/* The forward on the device's static workspace: the kernels and arguments of
 * forward(), with the workspace carved from the file-scope array of
 * float_mode.hip instead of this object's allocations, so that nothing is
 * allocated (the array's address is looked up on the device's first call).
 * Calls on one device share the array, so they must not overlap. */
void FloatNetwork::forwardStatic(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a,
		const float* b, const float* c, const float* d, const float* v0, const float* u0, const float* x,
		float* S, float* V, double* ckpt)
{
	if(T == 0) throw Error(NEMO_INVALID_INPUT, "the float-weight mode needs at least one step");
	if(K == 0) throw Error(NEMO_INVALID_INPUT, "checkpoint interval must be positive");
	HIP_CHECK(hipSetDevice(m_device));
	void* base = nullptr;
	size_t cap = 0;
	HIP_CHECK(nemo_float_scratch(&base, &cap));
	const size_t need = forwardWorkspaceBytes(T);
	if(need > cap) {
		throw Error(NEMO_INVALID_INPUT, "the allocation-free forward of " + std::to_string(T) + " steps needs "
				+ std::to_string(need) + " bytes of workspace; its static workspace holds " + std::to_string(cap));
	}
	forwardWith(carveWorkspace(base, T), st, T, K, w, a, b, c, d, v0, u0, x, S, V, ckpt);
}

/* Bytes of the forward's workspace for T steps, each part 256-byte aligned:
 * weights in reverse order (4 B per synapse), v and u in double (16 B per
 * neuron) and the spike bitmaps of every step. */
size_t FloatNetwork::forwardWorkspaceBytes(unsigned T) const
{
	auto up = [](size_t x) { return (x + 255) / 256 * 256; };
	return up(std::max<size_t>(1, m_s) * sizeof(float)) + 2 * up(m_n * sizeof(double))
		+ up(size_t(T) * m_pitch * sizeof(uint32_t));
}

FloatNetwork::ForwardWorkspace FloatNetwork::carveWorkspace(void* base, unsigned T) const
{
	auto up = [](size_t x) { return (x + 255) / 256 * 256; };
	unsigned char* p = static_cast<unsigned char*>(base);
	ForwardWorkspace ws;
	ws.wr = reinterpret_cast<float*>(p);
	p += up(std::max<size_t>(1, m_s) * sizeof(float));
	ws.v = reinterpret_cast<double*>(p);
	p += up(m_n * sizeof(double));
	ws.u = reinterpret_cast<double*>(p);
	p += up(m_n * sizeof(double));
	ws.bits = reinterpret_cast<uint32_t*>(p);
	(void) T;
	return ws;
}

/* The forward's kernels on a given workspace; launches only (nothing is
 * allocated, freed or waited for), so it may be captured in a graph. */
void FloatNetwork::forwardWith(const ForwardWorkspace& ws, hipStream_t st, unsigned T, unsigned K, const float* w,
		const float* a, const float* b, const float* c, const float* d, const float* v0, const float* u0,
		const float* x, float* S, float* V, double* ckpt)
{
	if(T == 0) throw Error(NEMO_INVALID_INPUT, "the float-weight mode needs at least one step");
	if(K == 0) throw Error(NEMO_INVALID_INPUT, "checkpoint interval must be positive");
	HIP_CHECK(hipSetDevice(m_device));
	HIP_CHECK(nemo_launch_gather(st, ws.wr, w, m_rcmUser.ptr, m_s));
	HIP_CHECK(nemo_launch_widen(st, ws.v, v0, m_n));
	HIP_CHECK(nemo_launch_widen(st, ws.u, u0, m_n));
	nemo_float_step_args f = {};
	f.n = unsigned(m_n);
	f.rcm_ptr = m_rcmPtr.ptr;
	f.rcm_sd = m_rcmSd.ptr;
	f.wr = ws.wr;
	f.pa = a;
	f.pb = b;
	f.pc = c;
	f.pd = d;
	f.x = x;
	f.S = S;
	f.V = V;
	f.v = ws.v;
	f.u = ws.u;
	f.store = 1;
	f.ckpt = ckpt;
	f.K = K;
	f.bits = ws.bits;
	f.pitch = m_pitch;
	f.dmax = m_D;
	f.stage = m_stage;
	for(unsigned t = 0; t < T; ++t) {
		f.t = t;
		HIP_CHECK(nemo_launch_float_forward_step(st, &f));
	}
}

/* Segments of K steps in reverse: replay the segment's forward steps from its
 * checkpoint (one launch), recording entry state and input, then run its
 * adjoint steps backwards. */
void FloatNetwork::backward(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a,
		const float* b, const float* c, const float* d, const float* x, const float* S, const double* ckpt,
		const float* gS, const float* gV, float* gw, float* ga, float* gb, float* gd, float* gx,
		float* gv0, float* gu0)
{
	prepare(T, K);
	HIP_CHECK(nemo_launch_gather(st, m_wr.ptr, w, m_rcmUser.ptr, m_s));
	HIP_CHECK(nemo_launch_fill_double(st, m_LV.ptr, 0.0, m_n));
	HIP_CHECK(nemo_launch_fill_double(st, m_LU.ptr, 0.0, m_n));
	HIP_CHECK(nemo_launch_fill_double(st, m_LI.ptr, 0.0, m_LI.n));
	HIP_CHECK(nemo_launch_fill(st, gw, 0.0f, m_s));
	HIP_CHECK(nemo_launch_fill(st, ga, 0.0f, m_n));
	HIP_CHECK(nemo_launch_fill(st, gb, 0.0f, m_n));
	HIP_CHECK(nemo_launch_fill(st, gd, 0.0f, m_n));
	const unsigned Kseg = std::min(K, T);
	const size_t segPlane = size_t(Kseg) * m_n;

	/* the replay reads the recorded spikes as firing histories */
	const unsigned rows = nemo_float_history_rows(T);
	if(size_t(rows) * m_n > m_hist.n) m_hist.allocate(size_t(rows) * m_n);
	HIP_CHECK(nemo_launch_pack_history(st, m_hist.ptr, rows, S, unsigned(m_n), T));
	nemo_float_replay_args f = {};
	f.n = unsigned(m_n);
	f.dmax = m_D;
	f.rcm_ptr = m_rcmPtr.ptr;
	f.rcm_sd = m_rcmSd.ptr;
	f.wr = m_wr.ptr;
	f.pa = a;
	f.pb = b;
	f.pc = c;
	f.pd = d;
	f.x = x;
	f.hist = m_hist.ptr;
	f.hist_rows = rows;
	f.seg_v = m_seg.ptr;
	f.seg_u = m_seg.ptr + segPlane;
	f.seg_I = m_seg.ptr + 2 * segPlane;

	nemo_float_back_args g = {};
	g.n = unsigned(m_n);
	g.T = T;
	g.ring = m_D + 1;
	g.fcm_ptr = m_fcmPtr.ptr;
	g.fcm_td = m_fcmTd.ptr;
	g.wf = w;
	g.pa = a;
	g.pb = b;
	g.pc = c;
	g.pd = d;
	g.S = S;
	g.gS = gS;
	g.gV = gV;
	g.seg_v = f.seg_v;
	g.seg_u = f.seg_u;
	g.seg_I = f.seg_I;
	g.LV = m_LV.ptr;
	g.LU = m_LU.ptr;
	g.LI = m_LI.ptr;
	g.gw = gw;
	g.ga = ga;
	g.gb = gb;
	g.gd = gd;
	g.gx = gx;

	const unsigned segments = (T + K - 1) / K;
	for(unsigned s = segments; s-- > 0;) {
		const unsigned t0 = s * K;
		const unsigned t1 = std::min(T, t0 + K);
		f.t0 = t0;
		f.t1 = t1;
		f.v0 = ckpt + size_t(s) * 2 * m_n;
		f.u0 = ckpt + size_t(s) * 2 * m_n + m_n;
		HIP_CHECK(nemo_launch_float_replay(st, &f));
		g.t0 = t0;
		for(unsigned t = t1; t-- > t0;) {
			g.t = t;
			HIP_CHECK(nemo_launch_float_backward_step(st, &g));
		}
	}
	if(gv0) HIP_CHECK(nemo_launch_narrow(st, gv0, m_LV.ptr, m_n));
	if(gu0) HIP_CHECK(nemo_launch_narrow(st, gu0, m_LU.ptr, m_n));
}

} // namespace nemo_hip
