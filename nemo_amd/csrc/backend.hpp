/*
 * Host side of the HIP backend: network construction, configuration, model
 * plugin loading, the connectivity builder and the simulation object behind
 * the C ABI in include/nemo_hip.h.
 *
 * Reference interfaces mirrored (NeMo v0.7.2):
 *   Network        src/nemo/Network.hpp:45-231, NetworkImpl.cpp:27-128
 *   Configuration  src/nemo/Configuration.hpp:70-195, ConfigurationImpl.cpp
 *   Simulation     src/nemo/Simulation.hpp:38-218, SimulationBackend.hpp:39-163
 *   NeuronType     src/nemo/NeuronType.cpp:26-158 (.ini metadata)
 *   Plugin         src/nemo/Plugin.cpp:31-195 (model shared libraries)
 */
#ifndef NEMO_AMD_BACKEND_HPP
#define NEMO_AMD_BACKEND_HPP

#include <chrono>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <algorithm>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "nemo_hip.h"
#include "nemo_hip_plugin.h"
#include "kernels.h"

namespace nemo_hip {

/* nemo::exception (src/nemo/exception.hpp:23-38): an error with a status code */
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void hipCheck(hipError_t err, const char* what);
#define HIP_CHECK(x) ::nemo_hip::hipCheck((x), #x)

/* Neuron model metadata from <Model>.ini (NeuronType.cpp:93-158) */
struct NeuronType {
	std::string name;
	unsigned nParam = 0;
	unsigned nState = 0;
	unsigned membranePotential = 0;
	unsigned history = 1;
	bool rngNormal = false;
	/* fields of the reverse matrix the model reads ([rcm]) */
	bool rcmSources = false, rcmDelays = false, rcmForward = false, rcmWeights = false;
	unsigned scratchPlanes = 0; /* [hip] scratch-planes: per-neuron device planes the plugin may use */
	std::string pluginDir;
};

NeuronType loadNeuronType(const std::string& name);
void addPluginPath(const std::string& dir);
std::vector<std::string> pluginPaths();

/* Network under construction (NetworkImpl) */
class Network {
public:
	unsigned addNeuronType(const std::string& name);
	void addNeurons(unsigned type, size_t n, const unsigned* idx, unsigned nargs, const float* args);
	void addSynapses(size_t n, const unsigned* src, const unsigned* tgt, const unsigned* delay,
			const float* weight, const unsigned char* plastic, synapse_id* ids);
	unsigned neuronCount() const;

	float getNeuronState(unsigned idx, unsigned var) const;
	float getNeuronParameter(unsigned idx, unsigned param) const;
	void setNeuron(unsigned idx, unsigned nargs, const float* args);
	void setNeuronState(unsigned idx, unsigned var, float val);
	void setNeuronParameter(unsigned idx, unsigned param, float val);

	struct Axon {
		std::vector<uint32_t> target;
		std::vector<uint8_t> delay;
		std::vector<float> weight;
		std::vector<uint8_t> plastic;
	};
	const Axon& axon(unsigned source) const;
	const std::vector<synapse_id>& getSynapsesFrom(unsigned source);

	struct Group {
		std::vector<std::vector<float> > param; /* [p][i] */
		std::vector<std::vector<float> > state; /* [s][i] */
		std::vector<unsigned> gidx;
	};
	std::vector<NeuronType> types;
	std::map<std::string, unsigned> typeIds;
	std::vector<Group> groups;
	std::unordered_map<unsigned, std::pair<unsigned, unsigned> > mapper; /* gidx -> (type, i) */
	std::map<unsigned, Axon> fcm; /* source -> synapses in insertion order */
	unsigned maxDelay = 0;
	size_t synapseCount = 0;

private:
	std::pair<unsigned, unsigned> address(unsigned idx) const;
	std::vector<synapse_id> m_queried;
};

/* Configuration (ConfigurationImpl) */
struct Configuration {
	bool stdp = false;
	std::vector<float> prefire, postfire;
	float minE = 0, maxE = 0, minI = 0, maxI = 0;
	unsigned fbits = 20; /* ConfigurationImpl.cpp:22 */
	backend_t backend = NEMO_BACKEND_HIP;
	int device = -1;
	bool writeOnlySynapses = false;
	bool logging = false;
	std::string description;

	void setStdpFunction(const std::vector<float>& pre, const std::vector<float>& post,
			float minE, float maxE, float minI, float maxI);
	void setHipBackend(int dev);
};

/* A loaded model plugin (Plugin.cpp:141-180) */
struct Plugin {
	void* handle = nullptr;
	hip_update_neurons_t* update = nullptr;
	hip_init_neurons_t* init = nullptr;
	explicit Plugin(const NeuronType& type);
	~Plugin();
	Plugin(const Plugin&) = delete;
	Plugin& operator=(const Plugin&) = delete;
};

/* Owning device allocation */
template<typename T>
struct DeviceArray {
	T* ptr = nullptr;
	size_t n = 0;
	DeviceArray() {}
	explicit DeviceArray(size_t count) { allocate(count); }
	~DeviceArray() { release(); }
	DeviceArray(const DeviceArray&) = delete;
	DeviceArray& operator=(const DeviceArray&) = delete;
	DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
	DeviceArray& operator=(DeviceArray&& o) noexcept {
		if(this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
		return *this;
	}
	void allocate(size_t count);
	void release();
	size_t bytes() const { return n * sizeof(T); }
};

/* Launch shape of the per-step kernels; any shape gives the same results. */
struct LaunchShape {
	unsigned deliverBlocks = 1;   /* delivery blocks for the delays 2..D (all delays if lateBlocks == 0) */
	unsigned lateBlocks = 0;      /* delivery blocks for the delay 1 (both sets <= NEMO_MAX_BIN_BLOCKS) */
	unsigned binCapacity = 4096;  /* events a delivery block stages in LDS */
	unsigned stdpMode = 2;        /* STDP accumulation: own launch on a graph branch joined before the
	                                 update (0), inside the update kernels (1), sharing the delivery
	                                 launch (2), own branch joined one step later (3), or in one
	                                 launch after each run of up to 64 steps (4) */
	unsigned stdpWords = 2;       /* firing-bitmap words (64 targets each) per STDP block */
	unsigned overlap = 1;         /* when split, deliver the delays 2..D of a step early: 1 on a graph
	                                 branch beside the previous update, 2 as extra blocks of the
	                                 previous update launch, 3 as in 2 with the previous update's
	                                 tiles pushing the delay-1 events too (one launch per step) */
};

/* The HIP simulation (cuda::Simulation's role, cuda/Simulation.cpp:91-380).
 *
 * A simulation advances one network, or a batch of independent networks in
 * the same launches: each network keeps its own neuron indices, RNG seeds
 * (as in its own simulation) and firing output; neuron types are shared by
 * name.  Single-network calls address network 0. */
class Simulation {
public:
	Simulation(const Network& net, const Configuration& conf);
	Simulation(const std::vector<const Network*>& nets, const Configuration& conf);
	~Simulation();

	unsigned networkCount() const { return m_nets; }

	/* SimulationBackend::step (SimulationBackend.cpp:61-77) */
	const std::vector<unsigned>& step(const unsigned* fstim, size_t nf,
			const unsigned* istimIdx, const float* istimVal, size_t ni);
	void applyStdp(float reward);
	/* Brian's propagate (Simulation.hpp:93-113): the firing of the next step
	 * is given (fired neuron indices, or one device flag per neuron in
	 * simulation order) instead of computed; the step's spikes are delivered
	 * and the synaptic input of the step after it is returned, E and I in
	 * simulation order (host arrays of neuronCount() floats, or device
	 * arrays valid until the next call).  Neuron state is not updated. */
	void propagate(const unsigned* fired, size_t nfired, float* E, float* I, size_t n);
	void propagateDevice(const uint32_t* flags, float** E, float** I);
	/* Where run() writes fired neurons: a caller's buffer while it has room,
	 * then a vector. */
	struct FiredSink {
		unsigned* buf = nullptr;
		size_t cap = 0, len = 0;
		std::vector<unsigned>* spill = nullptr;
		void push_back(unsigned v)
		{
			if(len < cap) buf[len++] = v;
			else spill->push_back(v);
		}
		void append(const unsigned* p, size_t n)
		{
			const size_t k = std::min(n, cap - len);
			std::copy(p, p + k, buf + len);
			len += k;
			spill->insert(spill->end(), p + k, p + n);
		}
		size_t size() const { return len + spill->size(); }
	};
	/* nsteps steps without stimulus; with readRaster, counts gets [step][network]
	 * firing counts and fired the fired neurons (by step, then network, then
	 * ascending simulation order) */
	void run(unsigned nsteps, unsigned stdpPeriod, float reward, bool readRaster,
			std::vector<unsigned>* counts, FiredSink* fired, uint64_t* events);
	/* capture the step graphs a run(nsteps, stdpPeriod) uses */
	void prepare(unsigned nsteps, unsigned stdpPeriod);
	void profile(unsigned nsteps, double avgUs[4], uint64_t counters[4]);

	LaunchShape launchShape() const { return m_shape; }
	void setLaunchShape(const LaunchShape& s);
	/* time candidate launch shapes over stepsPerTrial steps each, from a
	 * snapshot of the simulation state that is restored afterwards; keep the
	 * fastest */
	LaunchShape autotune(unsigned stepsPerTrial);

	float getNeuronState(unsigned idx, unsigned var, unsigned net = 0);
	float getNeuronParameter(unsigned idx, unsigned param, unsigned net = 0);
	float getMembranePotential(unsigned idx, unsigned net = 0);
	void setNeuron(unsigned idx, unsigned nargs, const float* args);
	void setNeuronState(unsigned idx, unsigned var, float val);
	void setNeuronParameter(unsigned idx, unsigned param, float val);
	/* vector forms: parameter (param) or state variable var of n neurons */
	void getNeuronValues(bool param, unsigned var, size_t n, const unsigned* idx, float* out);
	void setNeuronValues(bool param, unsigned var, size_t n, const unsigned* idx, const float* vals);
	void setNeurons(size_t n, const unsigned* idx, unsigned nargs, const float* args);

	const std::vector<synapse_id>& getSynapsesFrom(unsigned source);
	unsigned getSynapseTarget(synapse_id id);
	unsigned getSynapseDelay(synapse_id id);
	float getSynapseWeight(synapse_id id);
	unsigned char getSynapsePlastic(synapse_id id);

	unsigned long elapsedWallclock() const;
	unsigned long elapsedSimulation();
	void resetTimer();

	/* inspection */
	void setDebug(bool on);
	void getCurrents(float* E, float* I, size_t n);
	void localToGlobal(unsigned* g, size_t n) const;
	void getRecentFiring(uint64_t* h, size_t n);
	void getRngState(uint32_t* s, size_t n);
	void getNeuronStates(unsigned var, float* out, size_t n);
	void getFixedWeights(int32_t* w, size_t n);
	void getStdpAccumulators(int32_t* acc, size_t n);
	size_t neuronCount() const { return m_ncount; }
	size_t synapseCount() const { return m_nsynapses; }
	size_t plasticCount() const { return m_nplastic; }
	uint64_t eventsDelivered();

	const std::vector<unsigned>& firedOut() const { return m_firedOut; }

private:
	struct Group {
		NeuronType type;
		std::unique_ptr<Plugin> plugin;
		unsigned start = 0, count = 0;
		DeviceArray<float> param, state, scratch;
		DeviceArray<uint32_t> rng;
		std::vector<uint32_t> levelPtr;   /* update passes (nemo_hip_neuron_group.levels), host */
		DeviceArray<uint32_t> levelIdx;
		nemo_hip_neuron_group dev;
	};
	/* plane var of the group's current state (its history slot) */
	float* statePlane(const Group& g, unsigned var) const;
	void advanceSteps(uint64_t k);
	void checkStatus();

	int m_device = 0;
	hipStream_t m_stream = nullptr;
	hipStream_t m_stdpStream = nullptr;  /* parallel branch for STDP in its own launch */
	hipEvent_t m_stdpFork = nullptr, m_stdpDone = nullptr;
	hipEvent_t m_stdpLag[2] = {nullptr, nullptr}; /* mode 3: branch of step parity i done */
	hipStream_t m_earlyStream = nullptr; /* branch for the delays 2..D of the next step */
	hipStream_t m_copyStream = nullptr;  /* firing-bitmap readout of run(), beside the next steps */
	hipStream_t m_restStream = nullptr;  /* the rest of a fired list longer than expected */
	/* with several neuron models, every model's update after the first runs
	 * on a branch of its own, so a step waits for the slowest model */
	std::vector<hipStream_t> m_groupStreams;
	std::vector<hipEvent_t> m_groupDone;
	hipEvent_t m_groupFork = nullptr;
	double m_firedPerStep = 0.0;          /* fired neurons per step in the last run() with a raster */
	mutable std::vector<std::vector<unsigned> > m_decodePer; /* decodeFiringBatch's per-network lists */
	mutable std::vector<unsigned> m_decodeBuf;               /* decodeFiring's list of one step */
	bool m_decodeInOrder = false;         /* batch whose device order lists each network in turn, in order */
	bool m_binRanged = false;             /* updates read only their tiles' column ranges */
	hipEvent_t m_earlyFork = nullptr, m_earlyDone[2] = {nullptr, nullptr};
	/* run(): a chunk's steps are done (ready) and its firing is copied out
	 * (copied), per half of the readout buffers; created once, as creating
	 * them per run delays a short run's first work */
	hipEvent_t m_readoutReady[2] = {nullptr, nullptr}, m_readoutCopied[2] = {nullptr, nullptr};
	bool m_stdpPending = false;          /* the STDP branch has not been joined yet */
	unsigned m_nets = 1;   /* networks in the batch */
	size_t m_ncount = 0;   /* neurons */
	size_t m_npad = 0;     /* device index space: each type padded to whole tiles */
	unsigned m_tiles = 0;
	unsigned m_D = 1;
	unsigned m_fbits = 20;
	size_t m_nsynapses = 0, m_nplastic = 0, m_nonEmptyRows = 0;
	bool m_writeOnly = false;
	bool m_logging = false;

	std::vector<unsigned> m_l2g;          /* device index -> user index (~0 for padding) */
	std::vector<unsigned> m_wordBase;     /* per 64 device indices: u such that m_l2g[64w + j] = u + j
	                                         for all its neurons, or ~0 */
	std::vector<uint16_t> m_netOf;        /* device index -> network (batches only) */
	std::vector<uint32_t> m_simOf;        /* device index -> simulation index in its network
	                                         (batches whose device type order differs from a
	                                         network's own type order) */
	std::vector<char> m_reorder;          /* per network: device order != simulation order */
	std::vector<uint32_t> m_order;        /* simulation index (networks back to back) -> device index */
	std::unordered_map<uint64_t, unsigned> m_g2l; /* network << 32 | user index -> device index */
	std::vector<int> m_typeOfLocal;       /* device index -> group (-1 for padding) */
	std::vector<Group> m_groups;

	/* device state */
	DeviceArray<nemo_hip_clock> m_clock;
	DeviceArray<uint64_t> m_firedBits;    /* NEMO_STEP_RING firing bitmaps */
	size_t m_firedWords = 0;
	DeviceArray<int64_t> m_acc;           /* overflow accumulators */
	size_t m_accPitch = 0;
	DeviceArray<uint64_t> m_hist;     /* [2][N] firing history (NEMO_HIST_INDEX) */
	DeviceArray<uint64_t> m_stdpSnap; /* [N][2] history snapshots for STDP batches */
	DeviceArray<float> m_ext;
	DeviceArray<uint32_t> m_fstim;
	DeviceArray<float> m_debugCurrent;
	bool m_debug = false;
	DeviceArray<uint32_t> m_rowptr;
	DeviceArray<uint2> m_d1Rows;      /* [N] the delay-1 row of each device index: fcm [x, y) */
	DeviceArray<int2> m_fcm;
	DeviceArray<uint32_t> m_rcmPtr, m_rcmSd, m_rcmFwd;
	DeviceArray<uint32_t> m_rcmOf;    /* per FCM synapse: accumulator index, ~0 if static */
	/* all-synapse reverse matrix (models whose .ini asks for [rcm] fields) */
	DeviceArray<uint32_t> m_rcmAllPtr, m_rcmAllSd;
	DeviceArray<float> m_rcmAllW;
	DeviceArray<unsigned> m_status;   /* NEMO_HIP_STATUS_* bits set by the kernels */
	bool m_checkStatus = false;       /* some model may set status bits */
	DeviceArray<int32_t> m_rcmAcc;
	DeviceArray<uint32_t> m_stimIdx;
	DeviceArray<float> m_stimVal;
	DeviceArray<int2> m_bins;
	DeviceArray<uint32_t> m_binTable;
	DeviceArray<uint32_t> m_pairBounds, m_binRange;
	DeviceArray<uint2> m_windows;
	DeviceArray<unsigned long long> m_blockEvents, m_blockExamined;
	/* propagate: device index -> simulation index (~0 for padding), the
	 * given firing, and the input it produces */
	DeviceArray<uint32_t> m_devToSim, m_extFlags;
	DeviceArray<float> m_extE, m_extI;
	std::vector<uint2> m_srcWindow;       /* per source tile: target tiles [x, y) of its synapses */
	std::vector<uint64_t> m_pairWeight;   /* synapses per (source tile, delay) pair */

	/* STDP */
	bool m_stdp = false;
	int32_t m_minE = 0, m_maxE = 0, m_minI = 0, m_maxI = 0;
	uint64_t m_flushed = ~uint64_t(0); /* step whose STDP accumulation ran on its own */
	uint64_t m_base = 0;               /* the step base last set on the device (setBase) */

	/* propagate-kernel arguments and launch shape */
	nemo_propagate_args m_propagate;     /* common delivery arguments */
	nemo_propagate_args m_early, m_late;  /* the two delivery block sets */
	nemo_hip_stdp m_stdpArgs;            /* STDP accumulation arguments (rcm_ptr NULL: none) */
	LaunchShape m_shape;
	unsigned m_stdpBlocks = 0;
	size_t m_ldsBudget = 0;
	unsigned m_cuCount = 256;         /* compute units of the device */
	size_t m_ldsPerCu = 163840;       /* LDS bytes per compute unit */

	/* host side of synapse queries (network order = source ascending, then insertion) */
	std::map<uint64_t, std::pair<size_t, size_t> > m_synRange; /* network << 32 | gsrc -> [first,last) */
	std::vector<uint32_t> m_synAddr;   /* synapse -> fcm index */
	std::vector<uint32_t> m_synTarget; /* synapse -> user target index */
	std::vector<uint8_t> m_synDelay;
	std::vector<uint8_t> m_synPlastic;
	std::vector<int2> m_fcmHost;       /* cached copy of the fcm for weight queries */
	bool m_fcmHostValid = false;
	std::vector<synapse_id> m_queried;

	/* host staging */
	uint64_t m_hostStep = 0;           /* steps launched */
	uint64_t m_cycleOrigin = 0;        /* step of the last timer reset */
	/* cycle (step - m_cycleOrigin) of the last neuron update, when any ran:
	 * a model with state history H holds its current state in slot
	 * (m_lastCycle + 1) % H (cpu/Neurons.cpp:123), else in slot 0 */
	uint64_t m_lastCycle = 0;
	bool m_updated = false;
	static const unsigned SCALAR_SLOTS = 256;
	uint64_t* m_pinnedScalars = nullptr; /* ring of sources for host-to-device scalars */
	unsigned m_scalarNext = 0;
	uint64_t* m_pinnedBits = nullptr;  /* firing bitmaps read back */
	/* run()'s firing readout of a single network, compacted on the device:
	 * per half (chunks alternate), the steps' counts and fired lists */
	DeviceArray<uint32_t> m_l2gDev;       /* m_l2g on the device */
	DeviceArray<uint32_t> m_listDev;      /* [2][CHUNK * m_npad] fired lists */
	DeviceArray<uint32_t> m_listCountsDev; /* [2][CHUNK] */
	uint32_t* m_pinnedListCounts = nullptr; /* [2][CHUNK] */
	uint32_t* m_pinnedList[2] = {nullptr, nullptr};
	size_t m_pinnedListCap[2] = {0, 0};
	bool ensureListReadout();
	void waitEvent(hipEvent_t);
	void waitStream(hipStream_t);
	void growPinnedList(int half, size_t need, size_t keep);
	unsigned long long* m_pinnedCounters = nullptr; /* [2][m_blockEvents.n]: the event counters at the
	                                                   start and at the end of a run */
	std::vector<unsigned> m_firedOut;
	std::chrono::steady_clock::time_point m_wallStart;

	/* captured multi-step graphs, keyed by step count */
	std::map<unsigned, hipGraphExec_t> m_graphs;

	void activate() const;
	unsigned localIdx(unsigned g, unsigned net = 0) const;
	Group& groupOf(unsigned l);
	void construct(const std::vector<const Network*>& nets, const Configuration& conf);
	void buildNeurons(const std::vector<const Network*>& nets);
	void buildConnectivity(const std::vector<const Network*>& nets, const Configuration& conf);
	void buildStdp(const Configuration& conf);
	void allocateStepState();
	LaunchShape defaultShape() const;
	void applyShape(const LaunchShape& s);
	size_t ldsNeeded(const LaunchShape& s) const;
	void resetStdpBatches();
	void checkShape(const LaunchShape& s) const;
	nemo_hip_step_buffers stepBuffers(uint64_t offset, bool ext, bool fstim) const;
	void copyScalar(uint64_t* dst, uint64_t value);
	void setBase(uint64_t base);
	void enqueueSteps(uint64_t first, unsigned k, bool ext, bool fstim, hipEvent_t* ev = nullptr,
			bool forkPending = true, bool advanceBase = false);
	/* whether the graphs of steps move the device's step base on themselves */
	bool graphsAdvanceBase() const;
	void launchDelivery(const nemo_propagate_args& set, uint64_t offset, hipStream_t st, bool withStdp,
			hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
	void flushStdp();
	bool enqueueFlush();
	void joinStdp();
	template<typename Out> void decodeFiring(const uint64_t* bits, Out& out) const;
	template<typename Out> void decodeFiringBatch(const uint64_t* bits, Out& out, std::vector<unsigned>& counts) const;
	void readBitmaps(uint64_t t0, unsigned k, uint64_t* dst = nullptr, hipStream_t st = nullptr);
	uint64_t sumCounters(const DeviceArray<unsigned long long>& a);
	/* the graph of k steps (forkPending: it runs the pending STDP batch first) */
	hipGraphExec_t graphFor(unsigned k, bool forkPending = true);
	void clearGraphs();
	struct Snapshot;
	void takeSnapshot(Snapshot& s);
	void restoreSnapshot(const Snapshot& s);
	double timeSteps(unsigned nsteps);
	void syncFcmHost();
	size_t synapseIndex(synapse_id id) const;
};

/* A network prepared for the float-weight mode and its adjoint
 * (float_mode.hip).  One Izhikevich neuron type with indices 0..N-1;
 * weights, parameters, inputs and outputs are caller-owned device arrays
 * (neurons by index, synapses in network order).  Calls on one object must
 * not overlap (they share its workspace). */
class FloatNetwork {
public:
	FloatNetwork(const Network& net, int device);

	size_t neurons() const { return m_n; }
	size_t synapses() const { return m_s; }
	unsigned maxDelay() const { return m_D; }

	void forward(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a, const float* b,
			const float* c, const float* d, const float* v0, const float* u0, const float* x,
			float* S, float* V, double* ckpt);
	/* The device memory the forward uses beyond its arguments */
	struct ForwardWorkspace {
		float* wr;        /* weights in reverse order */
		double* v;        /* state */
		double* u;
		uint32_t* bits;   /* [T][pitch] spike bitmaps */
	};
	size_t forwardWorkspaceBytes(unsigned T) const;
	ForwardWorkspace carveWorkspace(void* base, unsigned T) const;
	void forwardWith(const ForwardWorkspace& ws, hipStream_t st, unsigned T, unsigned K, const float* w,
			const float* a, const float* b, const float* c, const float* d, const float* v0, const float* u0,
			const float* x, float* S, float* V, double* ckpt);
	/* forward on the device's static workspace (nemo_float_scratch): calls
	 * no allocator; fails if the workspace is too small for T steps */
	void forwardStatic(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a, const float* b,
			const float* c, const float* d, const float* v0, const float* u0, const float* x,
			float* S, float* V, double* ckpt);
	int device() const { return m_device; }
	void backward(hipStream_t st, unsigned T, unsigned K, const float* w, const float* a, const float* b,
			const float* c, const float* d, const float* x, const float* S, const double* ckpt,
			const float* gS, const float* gV, float* gw, float* ga, float* gb, float* gd, float* gx,
			float* gv0, float* gu0);

private:
	int m_device = 0;
	size_t m_n = 0, m_s = 0;
	unsigned m_D = 1;
	DeviceArray<uint32_t> m_fcmPtr, m_fcmTd;          /* network order is forward order */
	DeviceArray<uint32_t> m_rcmPtr, m_rcmSd, m_rcmUser;
	DeviceArray<float> m_wr;
	DeviceArray<double> m_v, m_u, m_seg;  /* state in double; segment: entry v, u and input I */
	DeviceArray<double> m_LV, m_LU, m_LI; /* adjoint state and recent input adjoints */
	unsigned m_segK = 0;
	DeviceArray<uint32_t> m_bits;         /* [T][m_pitch] spike bitmaps of the forward */
	DeviceArray<uint64_t> m_hist;         /* firing histories for the replay */
	unsigned m_pitch = 0, m_bitsT = 0;
	bool m_stage = false;

	void prepare(unsigned T, unsigned K);
};

} // namespace nemo_hip

#endif
