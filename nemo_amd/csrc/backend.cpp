/*
 * Host side of the HIP backend; see backend.hpp for the interfaces mirrored.
 */
#include "backend.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

#include <dlfcn.h>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace nemo_hip {

void hipCheck(hipError_t err, const char* what)
{
	if(err != hipSuccess) {
		int code = (err == hipErrorOutOfMemory) ? NEMO_CUDA_MEMORY_ERROR : NEMO_CUDA_INVOCATION_ERROR;
		throw Error(code, std::string("HIP error in ") + what + ": " + hipGetErrorString(err));
	}
}

template<typename T>
void DeviceArray<T>::allocate(size_t count)
{
	release();
	n = count;
	if(count) {
		HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
	}
}

template<typename T>
void DeviceArray<T>::release()
{
	if(ptr) {
		(void) hipFree(ptr);
		ptr = nullptr;
	}
	n = 0;
}

template struct DeviceArray<float>;
template struct DeviceArray<double>;
template struct DeviceArray<uint32_t>;
template struct DeviceArray<int32_t>;
template struct DeviceArray<int64_t>;
template struct DeviceArray<uint64_t>;
template struct DeviceArray<int2>;
template struct DeviceArray<nemo_hip_clock>;
template struct DeviceArray<unsigned long long>;
template struct DeviceArray<int4>;
template struct DeviceArray<uint2>;
template struct DeviceArray<unsigned char>;

/* ----------------------------------------------------- fixed point (Q11.20) */

/* fx_toFix, src/nemo/fixedpoint.cpp:17-28 */
static int32_t fxToFix(float f, unsigned fbits)
{
	if(std::abs(int(f)) >= (1 << (31 - fbits))) {
		std::ostringstream msg;
		msg << "Fixed-point overflow. Value " << f << " does not fit into fixed-point format Q"
			<< 31 - fbits << "." << fbits;
		throw Error(NEMO_INVALID_INPUT, msg.str());
	}
	return static_cast<int32_t>(f * float(1 << fbits));
}

static float fxToFloat(int32_t v, unsigned fbits)
{
	return float(v) / float(1 << fbits);
}

/* ------------------------------------------------------------ plugin paths */

static std::vector<std::string>& userPluginPaths()
{
	static std::vector<std::string> paths;
	return paths;
}

void addPluginPath(const std::string& dir)
{
	userPluginPaths().push_back(dir);
}

static std::string libraryDirectory()
{
	Dl_info info;
	if(dladdr(reinterpret_cast<void*>(&libraryDirectory), &info) && info.dli_fname) {
		std::string p(info.dli_fname);
		size_t slash = p.rfind('/');
		return slash == std::string::npos ? std::string(".") : p.substr(0, slash);
	}
	return ".";
}

/* Search order: directories added at run time, then NEMO_PLUGIN_PATH, then the
 * plugins directory installed beside the library (NeuronType.cpp:47-72). */
std::vector<std::string> pluginPaths()
{
	std::vector<std::string> paths = userPluginPaths();
	if(const char* env = std::getenv("NEMO_PLUGIN_PATH")) {
		std::stringstream ss(env);
		std::string item;
		while(std::getline(ss, item, ':')) {
			if(!item.empty()) paths.push_back(item);
		}
	}
	paths.push_back(libraryDirectory() + "/../plugins");
	return paths;
}

static bool fileExists(const std::string& p)
{
	std::ifstream f(p);
	return f.good();
}

static std::string trim(const std::string& s)
{
	size_t a = s.find_first_not_of(" \t\r\n");
	if(a == std::string::npos) return "";
	size_t b = s.find_last_not_of(" \t\r\n");
	return s.substr(a, b - a + 1);
}

/* boost::program_options' spellings of a bool (NeuronType.cpp reads the
 * .ini with program_options) */
static bool parseBool(const std::string& v, const std::string& key, const std::string& path)
{
	std::string l(v);
	std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return char(std::tolower(c)); });
	if(l == "true" || l == "1" || l == "yes" || l == "on") return true;
	if(l == "false" || l == "0" || l == "no" || l == "off") return false;
	throw Error(NEMO_INVALID_INPUT, "Error parsing neuron model configuration file " + path + ": invalid value '"
			+ v + "' for " + key);
}

static unsigned parseUnsigned(const std::string& v, const std::string& key, const std::string& path)
{
	size_t used = 0;
	unsigned long x = 0;
	try {
		x = std::stoul(v, &used);
	} catch(const std::exception&) {
		used = 0;
	}
	if(used == 0 || used != v.size() || v[0] == '-') {
		throw Error(NEMO_INVALID_INPUT, "Error parsing neuron model configuration file " + path
				+ ": invalid value '" + v + "' for " + key);
	}
	return unsigned(x);
}

/* <Model>.ini in NeMo's format (NeuronType.cpp:93-158): key=value lines,
 * [section] headers, '#' comments.  Required: parameters, state-variables,
 * membrane-potential and rng.normal; optional: history (1), rcm.sources,
 * rcm.delays, rcm.forward, rcm.weights (false) and the backend flags, of
 * which this backend needs backends.hip.  Other keys and sections -- the
 * [params:...] and [state_names] defaults NeMo's tools read -- are
 * ignored, as NeMo's parser does.  [hip] scratch-planes is this backend's
 * own. */
NeuronType loadNeuronType(const std::string& name)
{
	for(const std::string& dir : pluginPaths()) {
		std::string path = dir + "/" + name + ".ini";
		if(!fileExists(path)) continue;
		std::ifstream in(path);
		if(!in) {
			throw Error(NEMO_IO_ERROR, "Failed to open neuron model configuration file " + path);
		}
		NeuronType t;
		t.name = name;
		t.pluginDir = dir;
		std::string section, line;
		bool hip = false;
		bool haveParams = false, haveState = false, haveMp = false, haveRng = false;
		while(std::getline(in, line)) {
			size_t hash = line.find('#');
			if(hash != std::string::npos) line = line.substr(0, hash);
			line = trim(line);
			if(line.empty()) continue;
			if(line.front() == '[' && line.back() == ']') {
				section = trim(line.substr(1, line.size() - 2));
				continue;
			}
			size_t eq = line.find('=');
			if(eq == std::string::npos) continue;
			std::string key = trim(line.substr(0, eq));
			std::string val = trim(line.substr(eq + 1));
			if(section.empty()) {
				if(key == "parameters") { t.nParam = parseUnsigned(val, key, path); haveParams = true; }
				else if(key == "state-variables") { t.nState = parseUnsigned(val, key, path); haveState = true; }
				else if(key == "membrane-potential") { t.membranePotential = parseUnsigned(val, key, path); haveMp = true; }
				else if(key == "history") t.history = parseUnsigned(val, key, path);
			} else if(section == "rng") {
				if(key == "normal") { t.rngNormal = parseBool(val, "rng.normal", path); haveRng = true; }
			} else if(section == "rcm") {
				if(key == "sources") t.rcmSources = parseBool(val, "rcm.sources", path);
				else if(key == "delays") t.rcmDelays = parseBool(val, "rcm.delays", path);
				else if(key == "forward") t.rcmForward = parseBool(val, "rcm.forward", path);
				else if(key == "weights") t.rcmWeights = parseBool(val, "rcm.weights", path);
			} else if(section == "backends") {
				if(key == "hip") hip = parseBool(val, "backends.hip", path);
			} else if(section == "hip") {
				if(key == "scratch-planes") t.scratchPlanes = parseUnsigned(val, "hip.scratch-planes", path);
			}
		}
		const char* missing = !haveParams ? "parameters" : !haveState ? "state-variables"
			: !haveMp ? "membrane-potential" : !haveRng ? "rng.normal" : nullptr;
		if(missing) {
			throw Error(NEMO_INVALID_INPUT, std::string("Missing parameter '") + missing + "' in configuration file "
					+ path);
		}
		if(!hip) {
			throw Error(NEMO_API_UNSUPPORTED, "neuron type " + name + " has no HIP backend");
		}
		if(t.history < 1 || t.history > 64) {
			throw Error(NEMO_API_UNSUPPORTED, "neuron type " + name + ": state history length must be 1 to 64");
		}
		return t;
	}
	throw Error(NEMO_INVALID_INPUT, "could not find neuron type description " + name + ".ini in the plugin path");
}

Plugin::Plugin(const NeuronType& type)
{
	std::string path = type.pluginDir + "/hip/lib" + type.name + ".so";
	handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
	if(!handle) {
		throw Error(NEMO_DL_ERROR, std::string("failed to load neuron model plugin ") + path + ": " + dlerror());
	}
	update = reinterpret_cast<hip_update_neurons_t*>(dlsym(handle, "hip_update_neurons"));
	init = reinterpret_cast<hip_init_neurons_t*>(dlsym(handle, "hip_init_neurons"));
	if(!update || !init) {
		dlclose(handle);
		handle = nullptr;
		throw Error(NEMO_DL_ERROR, "plugin " + path + " does not export hip_update_neurons/hip_init_neurons");
	}
}

Plugin::~Plugin()
{
	/* Plugins stay mapped: their kernels' code objects are registered with the
	 * HIP runtime for the life of the process. */
}

/* ---------------------------------------------------------------- network */

unsigned Network::addNeuronType(const std::string& name)
{
	auto it = typeIds.find(name);
	if(it != typeIds.end()) return it->second;
	NeuronType t = loadNeuronType(name);
	unsigned id = unsigned(types.size());
	types.push_back(t);
	Group g;
	g.param.resize(t.nParam);
	g.state.resize(t.nState);
	groups.push_back(std::move(g));
	typeIds[name] = id;
	return id;
}

void Network::addNeurons(unsigned type, size_t n, const unsigned* idx, unsigned nargs, const float* args)
{
	if(type >= types.size()) {
		throw Error(NEMO_INVALID_INPUT, "Invalid neuron type id " + std::to_string(type));
	}
	const NeuronType& t = types[type];
	if(nargs != t.nParam + t.nState) {
		std::ostringstream msg;
		msg << "Unexpected number of parameters/state variables when adding neuron. Expected "
			<< (t.nParam + t.nState) << ", found " << nargs;
		throw Error(NEMO_INVALID_INPUT, msg.str());
	}
	Group& g = groups[type];
	for(size_t i = 0; i < n; ++i) {
		if(!mapper.emplace(idx[i], std::make_pair(type, unsigned(g.gidx.size()))).second) {
			throw Error(NEMO_INVALID_INPUT, "Duplicate neuron index " + std::to_string(idx[i]));
		}
		const float* a = args + i * nargs;
		for(unsigned p = 0; p < t.nParam; ++p) g.param[p].push_back(*a++);
		for(unsigned s = 0; s < t.nState; ++s) g.state[s].push_back(*a++);
		g.gidx.push_back(idx[i]);
	}
}

void Network::addSynapses(size_t n, const unsigned* src, const unsigned* tgt, const unsigned* delay,
		const float* weight, const unsigned char* plastic, synapse_id* ids)
{
	for(size_t i = 0; i < n; ++i) {
		if(delay[i] < 1) {
			std::ostringstream msg;
			msg << "Invalid delay (" << delay[i] << ") for synapse between " << src[i] << " and " << tgt[i];
			throw Error(NEMO_INVALID_INPUT, msg.str());
		}
		if(delay[i] > 255) {
			throw Error(NEMO_INVALID_INPUT, "Delay " + std::to_string(delay[i]) + " exceeds 255 ms");
		}
	}
	Axon* last = nullptr;
	unsigned lastSrc = 0;
	for(size_t i = 0; i < n; ++i) {
		if(!last || src[i] != lastSrc) {
			last = &fcm[src[i]];
			lastSrc = src[i];
		}
		Axon& ax = *last;
		if(ids) ids[i] = (synapse_id(src[i]) << 32) | synapse_id(ax.target.size());
		ax.target.push_back(tgt[i]);
		ax.delay.push_back(uint8_t(delay[i]));
		ax.weight.push_back(weight[i]);
		ax.plastic.push_back(plastic ? (plastic[i] != 0) : 0);
		maxDelay = std::max(maxDelay, delay[i]);
	}
	synapseCount += n;
}

unsigned Network::neuronCount() const
{
	size_t c = 0;
	for(const Group& g : groups) c += g.gidx.size();
	return unsigned(c);
}

std::pair<unsigned, unsigned> Network::address(unsigned idx) const
{
	auto it = mapper.find(idx);
	if(it == mapper.end()) {
		throw Error(NEMO_INVALID_INPUT, "Non-existing neuron index " + std::to_string(idx));
	}
	return it->second;
}

float Network::getNeuronState(unsigned idx, unsigned var) const
{
	auto a = address(idx);
	if(var >= types[a.first].nState) throw Error(NEMO_INVALID_INPUT, "Invalid state variable index " + std::to_string(var));
	return groups[a.first].state[var][a.second];
}

float Network::getNeuronParameter(unsigned idx, unsigned param) const
{
	auto a = address(idx);
	if(param >= types[a.first].nParam) throw Error(NEMO_INVALID_INPUT, "Invalid parameter index " + std::to_string(param));
	return groups[a.first].param[param][a.second];
}

void Network::setNeuron(unsigned idx, unsigned nargs, const float* args)
{
	auto a = address(idx);
	const NeuronType& t = types[a.first];
	if(nargs != t.nParam + t.nState) {
		throw Error(NEMO_INVALID_INPUT, "Unexpected number of parameters/state variables when modifying neuron");
	}
	for(unsigned p = 0; p < t.nParam; ++p) groups[a.first].param[p][a.second] = *args++;
	for(unsigned s = 0; s < t.nState; ++s) groups[a.first].state[s][a.second] = *args++;
}

void Network::setNeuronState(unsigned idx, unsigned var, float val)
{
	auto a = address(idx);
	if(var >= types[a.first].nState) throw Error(NEMO_INVALID_INPUT, "Invalid state variable index " + std::to_string(var));
	groups[a.first].state[var][a.second] = val;
}

void Network::setNeuronParameter(unsigned idx, unsigned param, float val)
{
	auto a = address(idx);
	if(param >= types[a.first].nParam) throw Error(NEMO_INVALID_INPUT, "Invalid parameter index " + std::to_string(param));
	groups[a.first].param[param][a.second] = val;
}

const Network::Axon& Network::axon(unsigned source) const
{
	auto it = fcm.find(source);
	if(it == fcm.end()) {
		throw Error(NEMO_INVALID_INPUT, "synapses of non-existing neuron (" + std::to_string(source) + ") requested");
	}
	return it->second;
}

const std::vector<synapse_id>& Network::getSynapsesFrom(unsigned source)
{
	m_queried.clear();
	auto it = fcm.find(source);
	if(it != fcm.end()) {
		for(size_t k = 0; k < it->second.target.size(); ++k) {
			m_queried.push_back((synapse_id(source) << 32) | synapse_id(k));
		}
	}
	return m_queried;
}

/* ---------------------------------------------------------- configuration */

void Configuration::setStdpFunction(const std::vector<float>& pre, const std::vector<float>& post,
		float minE_, float maxE_, float minI_, float maxI_)
{
	/* StdpFunction.cpp:12-48 */
	if(maxI_ > 0.0f || minI_ > 0.0f) {
		throw Error(NEMO_INVALID_INPUT, "STDP function should not have positive limits for inhibitory synapses");
	}
	if(maxE_ < 0.0f || minE_ < 0.0f) {
		throw Error(NEMO_INVALID_INPUT, "STDP function should not have negative limits for excitatory synapses");
	}
	if(maxE_ < minE_) {
		throw Error(NEMO_INVALID_INPUT, "STDP should have maximum excitatory weight >= minimum excitatory weight");
	}
	if(maxI_ > minI_) {
		throw Error(NEMO_INVALID_INPUT, "STDP should have abs(maximum inhibitory weight) >= abs(minimum inhibitory weight)");
	}
	if(pre.size() + post.size() > 64) {
		throw Error(NEMO_INVALID_INPUT, "size of STDP window too large");
	}
	stdp = true;
	prefire = pre;
	postfire = post;
	minE = minE_;
	maxE = maxE_;
	minI = minI_;
	maxI = maxI_;
}

void Configuration::setHipBackend(int dev)
{
	int count = 0;
	if(hipGetDeviceCount(&count) != hipSuccess || count == 0) {
		throw Error(NEMO_CUDA_ERROR, "no HIP device available");
	}
	if(dev >= count) {
		throw Error(NEMO_INVALID_INPUT, "invalid HIP device " + std::to_string(dev));
	}
	backend = NEMO_BACKEND_HIP;
	device = dev < 0 ? 0 : dev;
	hipDeviceProp_t prop;
	HIP_CHECK(hipGetDeviceProperties(&prop, device));
	description = std::string("HIP backend (") + prop.name + ", " + prop.gcnArchName + ")";
}

/* static LDS of an update kernel that carries propagate blocks, beside the
 * dynamic buffer both kinds of block share (block scans' wave totals) */
static const size_t COLAUNCH_LDS_MARGIN = 1024;

/* ------------------------------------------------------------- simulation */

static const char* stdpModeName(unsigned mode)
{
	return mode == 0 ? "own branch" : mode == 1 ? "in the updates" : mode == 2 ? "in the delivery launch"
			: mode == 3 ? "own branch, one step of slack" : "batched after each run of up to 64 steps";
}

Simulation::Simulation(const Network& net, const Configuration& conf)
{
	construct(std::vector<const Network*>(1, &net), conf);
}

Simulation::Simulation(const std::vector<const Network*>& nets, const Configuration& conf)
{
	construct(nets, conf);
}

void Simulation::construct(const std::vector<const Network*>& nets, const Configuration& conf)
{
	if(conf.backend != NEMO_BACKEND_HIP) {
		throw Error(NEMO_API_UNSUPPORTED, "this library provides the HIP backend only; select it with nemo_set_hip_backend");
	}
	if(nets.empty() || nets.size() > 65535) {
		throw Error(NEMO_INVALID_INPUT, "a batch holds 1 to 65535 networks");
	}
	for(const Network* net : nets) {
		/* nemo.cpp:111-117 */
		if(!net || net->neuronCount() == 0) {
			throw Error(NEMO_INVALID_INPUT, "Cannot create simulation from empty network");
		}
		if(conf.stdp && net->maxDelay + conf.postfire.size() > 64) {
			throw Error(NEMO_INVALID_INPUT, "Max delay + postfire part of STDP function must be <= 64");
		}
		if(net->maxDelay > NEMO_MAX_DELAY) {
			throw Error(NEMO_INVALID_INPUT, "The network has synapses with delay " + std::to_string(net->maxDelay)
					+ "ms. The HIP backend supports a maximum of 64 ms");
		}
	}
	m_nets = unsigned(nets.size());
	m_device = conf.device < 0 ? 0 : conf.device;
	m_fbits = conf.fbits;
	m_writeOnly = conf.writeOnlySynapses;
	m_logging = conf.logging;
	activate();
	HIP_CHECK(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking));
	HIP_CHECK(hipStreamCreateWithFlags(&m_stdpStream, hipStreamNonBlocking));
	HIP_CHECK(hipEventCreateWithFlags(&m_stdpFork, hipEventDisableTiming));
	HIP_CHECK(hipEventCreateWithFlags(&m_stdpDone, hipEventDisableTiming));
	HIP_CHECK(hipEventCreateWithFlags(&m_stdpLag[0], hipEventDisableTiming));
	HIP_CHECK(hipEventCreateWithFlags(&m_stdpLag[1], hipEventDisableTiming));
	HIP_CHECK(hipStreamCreateWithFlags(&m_earlyStream, hipStreamNonBlocking));
	HIP_CHECK(hipStreamCreateWithFlags(&m_copyStream, hipStreamNonBlocking));
	HIP_CHECK(hipStreamCreateWithFlags(&m_restStream, hipStreamNonBlocking));
	HIP_CHECK(hipEventCreateWithFlags(&m_earlyFork, hipEventDisableTiming));
	HIP_CHECK(hipEventCreateWithFlags(&m_earlyDone[0], hipEventDisableTiming));
	HIP_CHECK(hipEventCreateWithFlags(&m_earlyDone[1], hipEventDisableTiming));
	for(int i = 0; i < 2; ++i) {
		HIP_CHECK(hipEventCreateWithFlags(&m_readoutReady[i], hipEventDisableTiming));
		HIP_CHECK(hipEventCreateWithFlags(&m_readoutCopied[i], hipEventDisableTiming));
	}
	HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_pinnedScalars), SCALAR_SLOTS * sizeof(uint64_t),
				hipHostMallocDefault));

	buildNeurons(nets);
	/* single-type batches (and any whose networks occupy increasing device
	 * ranges in their own order) decode their fired lists in one pass */
	m_decodeInOrder = m_nets > 1 && std::find(m_reorder.begin(), m_reorder.end(), 1) == m_reorder.end();
	for(size_t l = 0, prev = 0; m_decodeInOrder && l < m_npad; ++l) {
		if(m_netOf[l] == uint16_t(0xffff)) continue;
		if(m_netOf[l] < prev) m_decodeInOrder = false;
		prev = m_netOf[l];
	}
	buildConnectivity(nets, conf);
	buildStdp(conf);
	allocateStepState();
	applyShape(defaultShape());

	size_t models = 0;
	for(const Group& g : m_groups) models += g.dev.count ? 1 : 0;
	if(models > 1) {
		HIP_CHECK(hipEventCreateWithFlags(&m_groupFork, hipEventDisableTiming));
		m_groupStreams.resize(models - 1, nullptr);
		m_groupDone.resize(models - 1, nullptr);
		for(size_t q = 0; q + 1 < models; ++q) {
			HIP_CHECK(hipStreamCreateWithFlags(&m_groupStreams[q], hipStreamNonBlocking));
			HIP_CHECK(hipEventCreateWithFlags(&m_groupDone[q], hipEventDisableTiming));
		}
	}

	nemo_hip_step_buffers b = stepBuffers(0, false, false);
	for(Group& g : m_groups) {
		HIP_CHECK(g.plugin->init(m_stream, &g.dev, &b));
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	if(m_logging) {
		std::printf("nemo-hip: %u network(s), %zu neurons (%u tiles), %zu synapses (%zu plastic), max delay %u, "
				"Q%u.%u, delivery %u + %u blocks x %u events, STDP %u blocks (%s)\n",
				m_nets, m_ncount, m_tiles, m_nsynapses, m_nplastic, m_D, 31 - m_fbits, m_fbits,
				m_shape.deliverBlocks, m_shape.lateBlocks, m_shape.binCapacity, m_stdpBlocks,
				stdpModeName(m_shape.stdpMode));
	}
	resetTimer();
}

Simulation::~Simulation()
{
	(void) hipSetDevice(m_device);
	if(m_stream) (void) hipStreamSynchronize(m_stream);
	if(m_stdpStream) (void) hipStreamSynchronize(m_stdpStream);
	if(m_earlyStream) (void) hipStreamSynchronize(m_earlyStream);
	if(m_copyStream) (void) hipStreamSynchronize(m_copyStream);
	if(m_restStream) (void) hipStreamSynchronize(m_restStream);
	for(hipStream_t q : m_groupStreams) {
		if(q) (void) hipStreamSynchronize(q);
	}
	clearGraphs();
	if(m_copyStream) (void) hipStreamDestroy(m_copyStream);
	if(m_restStream) (void) hipStreamDestroy(m_restStream);
	for(hipStream_t q : m_groupStreams) {
		if(q) (void) hipStreamDestroy(q);
	}
	for(hipEvent_t e : m_groupDone) {
		if(e) (void) hipEventDestroy(e);
	}
	if(m_groupFork) (void) hipEventDestroy(m_groupFork);
	for(hipEvent_t e : {m_readoutReady[0], m_readoutReady[1], m_readoutCopied[0], m_readoutCopied[1]}) {
		if(e) (void) hipEventDestroy(e);
	}
	if(m_earlyStream) {
		(void) hipStreamDestroy(m_earlyStream);
		(void) hipEventDestroy(m_earlyFork);
		(void) hipEventDestroy(m_earlyDone[0]);
		(void) hipEventDestroy(m_earlyDone[1]);
	}
	if(m_stream) (void) hipStreamDestroy(m_stream);
	if(m_stdpStream) {
		(void) hipStreamDestroy(m_stdpStream);
		(void) hipEventDestroy(m_stdpFork);
		(void) hipEventDestroy(m_stdpDone);
		(void) hipEventDestroy(m_stdpLag[0]);
		(void) hipEventDestroy(m_stdpLag[1]);
	}
	if(m_pinnedScalars) (void) hipHostFree(m_pinnedScalars);
	if(m_pinnedBits) (void) hipHostFree(m_pinnedBits);
	if(m_pinnedCounters) (void) hipHostFree(m_pinnedCounters);
	if(m_pinnedListCounts) (void) hipHostFree(m_pinnedListCounts);
	for(uint32_t* p : m_pinnedList) {
		if(p) (void) hipHostFree(p);
	}
}

void Simulation::activate() const
{
	HIP_CHECK(hipSetDevice(m_device));
}

static inline uint64_t netKey(unsigned net, unsigned g)
{
	return (uint64_t(net) << 32) | g;
}

unsigned Simulation::localIdx(unsigned g, unsigned net) const
{
	if(net >= m_nets) {
		throw Error(NEMO_INVALID_INPUT, "Non-existing network " + std::to_string(net) + " in batch");
	}
	auto it = m_g2l.find(netKey(net, g));
	if(it == m_g2l.end()) {
		throw Error(NEMO_INVALID_INPUT, "Non-existing neuron index " + std::to_string(g));
	}
	return it->second;
}

Simulation::Group& Simulation::groupOf(unsigned l)
{
	return m_groups.at(m_typeOfLocal.at(l));
}

/* Simulation order of a network: type by type (addNeuronType order),
 * insertion order within a type (cpu/Simulation.cpp:60-77, cpu/Neurons.cpp:
 * 27-40).  On the device the types follow each other (by name, in order of
 * first appearance in the batch), each starting at a multiple of NEMO_TILE so
 * that one update block owns one tile of one type; within a type the
 * networks of a batch follow each other without padding.  RNG seeds: mt19937
 * default seed, four words >> 1 per simulation index of the neuron in its own
 * network (RNG.cpp:38-63, seeded by the CPU backend's local index,
 * cpu/Neurons.cpp:41), so that a network's random streams do not depend on
 * the batch it is in. */
void Simulation::buildNeurons(const std::vector<const Network*>& nets)
{
	std::vector<NeuronType> types;
	std::map<std::string, unsigned> slotOf;
	size_t maxNet = 0;
	m_ncount = 0;
	for(const Network* net : nets) {
		for(unsigned t = 0; t < net->types.size(); ++t) {
			if(net->groups[t].gidx.empty()) continue;
			if(slotOf.emplace(net->types[t].name, unsigned(types.size())).second) {
				types.push_back(net->types[t]);
			}
		}
		m_ncount += net->neuronCount();
		maxNet = std::max<size_t>(maxNet, net->neuronCount());
	}
	std::vector<size_t> typeCount(types.size(), 0);
	for(const Network* net : nets) {
		for(unsigned t = 0; t < net->types.size(); ++t) {
			if(!net->groups[t].gidx.empty()) typeCount[slotOf[net->types[t].name]] += net->groups[t].gidx.size();
		}
	}
	size_t npad = 0;
	for(size_t c : typeCount) npad += (c + NEMO_TILE - 1) / NEMO_TILE * NEMO_TILE;
	if(npad >= (size_t(1) << 26)) {
		throw Error(NEMO_API_UNSUPPORTED, "the HIP backend supports at most 2^26 neurons per simulation");
	}
	m_npad = npad;
	m_tiles = unsigned(npad / NEMO_TILE);
	m_l2g.assign(m_npad, ~0u);
	if(m_nets > 1) m_netOf.assign(m_npad, uint16_t(0xffff));
	/* a network whose types appear in another order than the device's must
	 * have its fired neurons reordered to its own simulation order */
	m_reorder.assign(nets.size(), 0);
	for(size_t k = 0; k < nets.size(); ++k) {
		unsigned last = 0;
		for(unsigned t = 0; t < nets[k]->types.size(); ++t) {
			if(nets[k]->groups[t].gidx.empty()) continue;
			const unsigned slot = slotOf[nets[k]->types[t].name];
			if(slot < last) m_reorder[k] = 1;
			last = slot;
		}
	}
	if(std::find(m_reorder.begin(), m_reorder.end(), 1) != m_reorder.end()) m_simOf.assign(m_npad, 0);
	m_typeOfLocal.assign(m_npad, -1);
	m_order.resize(m_ncount);
	m_g2l.reserve(m_ncount * 2);

	std::vector<uint32_t> seeds(4 * maxNet);
	{
		std::mt19937 mt;
		for(size_t i = 0; i < seeds.size(); ++i) seeds[i] = uint32_t(mt()) >> 1;
	}
	std::vector<size_t> netSimBase(nets.size() + 1, 0);
	for(size_t k = 0; k < nets.size(); ++k) netSimBase[k + 1] = netSimBase[k] + nets[k]->neuronCount();

	unsigned start = 0; /* device index */
	for(unsigned ds = 0; ds < types.size(); ++ds) {
		m_groups.emplace_back();
		Group& g = m_groups.back();
		g.type = types[ds];
		g.plugin.reset(new Plugin(g.type));
		g.start = start;
		g.count = unsigned(typeCount[ds]);
		const size_t pitch = g.count;
		std::vector<float> param(std::max<size_t>(1, g.type.nParam * pitch));
		/* slot 0 of the state history holds the initial state */
		std::vector<float> state(std::max<size_t>(1, size_t(g.type.history) * g.type.nState * pitch));
		std::vector<uint32_t> rng(4 * size_t(g.count));
		size_t pos = 0;
		for(unsigned k = 0; k < nets.size(); ++k) {
			const Network& net = *nets[k];
			auto tid = net.typeIds.find(g.type.name);
			if(tid == net.typeIds.end()) continue;
			const unsigned t = tid->second;
			const Network::Group& ng = net.groups[t];
			const size_t n = ng.gidx.size();
			if(n == 0) continue;
			size_t simBase = 0; /* simulation index of the type's first neuron in its own network */
			for(unsigned u = 0; u < t; ++u) simBase += net.groups[u].gidx.size();
			for(size_t i = 0; i < n; ++i) {
				const unsigned l = unsigned(start + pos + i);
				if(!m_g2l.emplace(netKey(k, ng.gidx[i]), l).second) {
					throw Error(NEMO_INVALID_INPUT, "Duplicate neuron index " + std::to_string(ng.gidx[i]));
				}
				m_l2g[l] = ng.gidx[i];
				if(m_nets > 1) m_netOf[l] = uint16_t(k);
				if(!m_simOf.empty()) m_simOf[l] = uint32_t(simBase + i);
				m_typeOfLocal[l] = int(m_groups.size() - 1);
				m_order[netSimBase[k] + simBase + i] = l;
			}
			std::copy(seeds.begin() + 4 * simBase, seeds.begin() + 4 * (simBase + n), rng.begin() + 4 * pos);
			for(unsigned p = 0; p < g.type.nParam; ++p) {
				std::copy(ng.param[p].begin(), ng.param[p].end(), param.begin() + p * pitch + pos);
			}
			for(unsigned v = 0; v < g.type.nState; ++v) {
				std::copy(ng.state[v].begin(), ng.state[v].end(), state.begin() + v * pitch + pos);
			}
			pos += n;
		}
		g.param.allocate(param.size());
		g.state.allocate(state.size());
		g.rng.allocate(rng.size());
		HIP_CHECK(hipMemcpy(g.param.ptr, param.data(), g.param.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(g.state.ptr, state.data(), g.state.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(g.rng.ptr, rng.data(), g.rng.bytes(), hipMemcpyHostToDevice));
		g.dev.start = g.start;
		g.dev.count = g.count;
		g.dev.pitch = pitch;
		g.dev.param = g.param.ptr;
		g.dev.state = g.state.ptr;
		g.dev.rng = g.rng.ptr;
		g.dev.scratch = nullptr;
		g.dev.history = g.type.history;
		g.dev.levels = 1;
		g.dev.level_ptr = nullptr;
		g.dev.level_idx = nullptr;
		if(g.type.scratchPlanes && g.count) {
			g.scratch.allocate(size_t(g.type.scratchPlanes) * pitch);
			g.dev.scratch = g.scratch.ptr;
		}
		start += (g.count + NEMO_TILE - 1) / NEMO_TILE * NEMO_TILE;
	}
	/* words of 64 device indices whose neurons have consecutive user indices
	 * (padding aside): the readout decodes them without the lookup table,
	 * whose scattered reads miss the host caches when a run is short */
	m_wordBase.assign(m_npad / 64, ~0u);
	for(size_t w = 0; w < m_npad / 64; ++w) {
		bool linear = true, any = false;
		unsigned base = 0;
		for(unsigned j = 0; j < 64 && linear; ++j) {
			const unsigned u = m_l2g[w * 64 + j];
			if(u == ~0u) continue;
			if(!any) {
				any = u >= j;
				linear = any;
				base = u - j;
			} else {
				linear = u == base + j;
			}
		}
		if(linear && any) m_wordBase[w] = base;
	}
}

/* The forward matrix: for every (source, delay) a contiguous row of
 * {target, Q11.20 weight}, rows in (source, delay) order, synapses within a
 * row in insertion order (ConnectivityMatrix.cpp:65-176).  The reverse matrix
 * holds, per target, its plastic incoming synapses (construction/RCM.cpp:
 * 112-178): packed source/delay and the forward index of the weight. */
void Simulation::buildConnectivity(const std::vector<const Network*>& nets, const Configuration& conf)
{
	m_D = 1;
	size_t total = 0;
	for(const Network* net : nets) {
		m_D = std::max(m_D, net->maxDelay);
		total += net->synapseCount;
	}
	const unsigned D = m_D;
	const size_t N = m_npad;
	if(total >= (size_t(1) << 32)) {
		throw Error(NEMO_API_UNSUPPORTED, "the HIP backend supports fewer than 2^32 synapses per simulation");
	}

	/* sources of all networks, each network's in ascending user index */
	struct Src { unsigned net; unsigned g; unsigned l; const Network::Axon* ax; size_t seq; };
	std::vector<Src> srcs;
	size_t seq = 0;
	for(unsigned k = 0; k < nets.size(); ++k) {
		for(auto& kv : nets[k]->fcm) {
			srcs.push_back(Src{k, kv.first, localIdx(kv.first, k), &kv.second, seq});
			seq += kv.second.target.size();
		}
	}
	m_nsynapses = seq;
	const long ns = long(srcs.size());

	/* device target index of every synapse, checked (verifySynapseTerminals) */
	std::vector<uint32_t> ltgt(m_nsynapses);
	std::vector<uint2> srcWin(srcs.size());
	int bad = 0;
	unsigned badIdx = 0;
#pragma omp parallel for schedule(dynamic, 64)
	for(long i = 0; i < ns; ++i) {
		const Src& s = srcs[i];
		uint32_t lo = ~0u, hi = 0;
		for(size_t k = 0; k < s.ax->target.size(); ++k) {
			auto it = m_g2l.find(netKey(s.net, s.ax->target[k]));
			if(it == m_g2l.end()) {
#pragma omp critical
				{ bad = 1; badIdx = s.ax->target[k]; }
				ltgt[s.seq + k] = 0;
			} else {
				ltgt[s.seq + k] = it->second;
				lo = std::min(lo, it->second / NEMO_TILE);
				hi = std::max(hi, it->second / NEMO_TILE + 1);
			}
		}
		srcWin[i] = make_uint2(lo, hi);
	}
	if(bad) {
		throw Error(NEMO_INVALID_INPUT, "Invalid synapse target neuron " + std::to_string(badIdx));
	}
	/* target tiles reached from each source tile */
	m_srcWindow.assign(m_tiles, make_uint2(~0u, 0));
	for(long i = 0; i < ns; ++i) {
		if(srcWin[i].y == 0) continue;
		uint2& w = m_srcWindow[srcs[i].l / NEMO_TILE];
		w.x = std::min(w.x, srcWin[i].x);
		w.y = std::max(w.y, srcWin[i].y);
	}

	/* row lengths and offsets */
	std::vector<uint32_t> rowptr(N * D + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
	for(long i = 0; i < ns; ++i) {
		const Src& s = srcs[i];
		uint32_t* cnt = &rowptr[size_t(s.l) * D + 1];
		for(uint8_t d : s.ax->delay) cnt[d - 1] += 1;
	}
	for(size_t r = 0; r < N * D; ++r) rowptr[r + 1] += rowptr[r];

	/* fill rows (stable within a source) */
	std::vector<int2> fcm(std::max<size_t>(1, m_nsynapses));
	m_synAddr.resize(m_nsynapses);
	if(!m_writeOnly) {
		m_synTarget.resize(m_nsynapses);
		m_synDelay.resize(m_nsynapses);
		m_synPlastic.resize(m_nsynapses);
	}
	int overflow = 0;
	float overflowValue = 0;
#pragma omp parallel for schedule(dynamic, 64)
	for(long i = 0; i < ns; ++i) {
		const Src& s = srcs[i];
		uint32_t cursor[NEMO_MAX_DELAY];
		std::copy(&rowptr[size_t(s.l) * D], &rowptr[size_t(s.l) * D] + D, cursor);
		for(size_t k = 0; k < s.ax->target.size(); ++k) {
			const unsigned d = s.ax->delay[k];
			const uint32_t pos = cursor[d - 1]++;
			const float w = s.ax->weight[k];
			int32_t wf = 0;
			if(std::abs(int(w)) >= (1 << (31 - m_fbits))) {
#pragma omp critical
				{ overflow = 1; overflowValue = w; }
			} else {
				wf = static_cast<int32_t>(w * float(1 << m_fbits));
			}
			fcm[pos].x = int(ltgt[s.seq + k]);
			fcm[pos].y = wf;
			m_synAddr[s.seq + k] = pos;
			if(!m_writeOnly) {
				m_synTarget[s.seq + k] = s.ax->target[k];
				m_synDelay[s.seq + k] = uint8_t(d);
				m_synPlastic[s.seq + k] = s.ax->plastic[k];
			}
		}
	}
	if(overflow) {
		fxToFix(overflowValue, m_fbits); /* throws with the reference message */
	}
	for(const Src& s : srcs) {
		m_synRange[netKey(s.net, s.g)] = std::make_pair(s.seq, s.seq + s.ax->target.size());
	}
	/* synapses per (source tile, delay) pair: the delivery work the blocks share */
	m_pairWeight.assign(size_t(m_tiles) * D, 0);
	for(size_t n = 0; n < N; ++n) {
		uint64_t* w = &m_pairWeight[(n / NEMO_TILE) * D];
		for(unsigned d = 0; d < D; ++d) w[d] += rowptr[n * D + d + 1] - rowptr[n * D + d];
	}

	m_rowptr.allocate(rowptr.size());
	HIP_CHECK(hipMemcpy(m_rowptr.ptr, rowptr.data(), m_rowptr.bytes(), hipMemcpyHostToDevice));
	{
		/* the delay-1 rows by themselves, for updates that push them */
		std::vector<uint2> d1(N);
		for(size_t n = 0; n < N; ++n) d1[n] = make_uint2(rowptr[n * D], rowptr[n * D + 1]);
		m_d1Rows.allocate(N);
		HIP_CHECK(hipMemcpy(m_d1Rows.ptr, d1.data(), m_d1Rows.bytes(), hipMemcpyHostToDevice));
	}
	m_fcm.allocate(fcm.size());
	HIP_CHECK(hipMemcpy(m_fcm.ptr, fcm.data(), m_fcm.bytes(), hipMemcpyHostToDevice));
	size_t rows = 0;
	for(size_t r = 0; r < N * D; ++r) rows += rowptr[r + 1] != rowptr[r];
	m_nonEmptyRows = rows;
	if(!m_writeOnly) {
		m_fcmHost.swap(fcm);
		m_fcmHostValid = true;
	}

	/* reverse matrix (plastic synapses only, STDP enabled) */
	m_stdp = conf.stdp;
	std::vector<uint32_t> rcmPtr(N + 1, 0);
	std::vector<uint32_t> rcmSd, rcmFwd;
	if(m_stdp && m_writeOnly) {
		throw Error(NEMO_API_UNSUPPORTED, "STDP with write-only synapses is not supported by the HIP backend");
	}
	if(m_stdp) {
		/* per-thread counts over contiguous source ranges give each thread a
		 * private, deterministic slot range in every target's row */
		int T = 1;
#ifdef _OPENMP
		T = std::max(1, omp_get_max_threads());
#endif
		const long chunk = (ns + T - 1) / T;
		std::vector<std::vector<uint32_t> > cnt(T, std::vector<uint32_t>(N, 0));
#pragma omp parallel for num_threads(T) schedule(static, 1)
		for(int t = 0; t < T; ++t) {
			for(long i = t * chunk; i < std::min(ns, (t + 1) * chunk); ++i) {
				const Src& s = srcs[i];
				for(size_t k = 0; k < s.ax->target.size(); ++k) {
					if(s.ax->plastic[k]) cnt[t][ltgt[s.seq + k]] += 1;
				}
			}
		}
		for(size_t n = 0; n < N; ++n) {
			uint32_t run = rcmPtr[n];
			for(int t = 0; t < T; ++t) {
				uint32_t c = cnt[t][n];
				cnt[t][n] = run;
				run += c;
			}
			rcmPtr[n + 1] = run;
		}
		m_nplastic = rcmPtr[N];
		rcmSd.resize(std::max<size_t>(1, m_nplastic));
		rcmFwd.resize(std::max<size_t>(1, m_nplastic));
#pragma omp parallel for num_threads(T) schedule(static, 1)
		for(int t = 0; t < T; ++t) {
			for(long i = t * chunk; i < std::min(ns, (t + 1) * chunk); ++i) {
				const Src& s = srcs[i];
				for(size_t k = 0; k < s.ax->target.size(); ++k) {
					if(!s.ax->plastic[k]) continue;
					const uint32_t tg = ltgt[s.seq + k];
					const uint32_t pos = cnt[t][tg]++;
					rcmSd[pos] = uint32_t(s.l) | (uint32_t(s.ax->delay[k] - 1) << 26);
					rcmFwd[pos] = m_synAddr[s.seq + k];
				}
			}
		}
	}
	/* all-synapse reverse matrix, for models that read their incoming
	 * synapses' sources, delays or weights (construction/RCM.cpp:76-98):
	 * per target, sources in ascending order and each source's synapses in
	 * insertion order, as the reference's CPU backend adds them */
	bool rcmAll = false;
	for(const Group& g : m_groups) {
		rcmAll = rcmAll || g.type.rcmSources || g.type.rcmDelays || g.type.rcmWeights || g.type.rcmForward;
	}
	if(rcmAll) {
		if(m_stdp) {
			throw Error(NEMO_API_UNSUPPORTED, "The current version does not support a mix of STDP with neuron types "
					"which require reverse connectivity for other purposes");
		}
		std::vector<uint32_t> ptr(N + 1, 0);
		for(const Src& s : srcs) {
			for(size_t k = 0; k < s.ax->target.size(); ++k) ptr[ltgt[s.seq + k] + 1] += 1;
		}
		for(size_t n = 0; n < N; ++n) ptr[n + 1] += ptr[n];
		std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
		std::vector<uint32_t> sd(std::max<size_t>(1, m_nsynapses));
		std::vector<float> w(std::max<size_t>(1, m_nsynapses));
		for(const Src& s : srcs) {
			for(size_t k = 0; k < s.ax->target.size(); ++k) {
				const uint32_t tg = ltgt[s.seq + k];
				const int tgroup = m_typeOfLocal[tg];
				const Group& g = m_groups[size_t(tgroup)];
				const bool reads = g.type.rcmSources || g.type.rcmDelays || g.type.rcmWeights || g.type.rcmForward;
				if(reads && m_typeOfLocal[s.l] != tgroup) {
					/* the models that read the matrix index their own state by
					 * the source (cpu/plugins/Kuramoto.cpp:52-78) */
					throw Error(NEMO_API_UNSUPPORTED, "synapse from neuron " + std::to_string(s.g) + " to a neuron of type "
							+ g.type.name + ", whose model reads its sources' state: sources must be of the same type");
				}
				const uint32_t pos = fill[tg]++;
				sd[pos] = uint32_t(s.l) | (uint32_t(s.ax->delay[k] - 1) << 26);
				w[pos] = s.ax->weight[k];
			}
		}
		/* update passes of the models with a state history that read the
		 * slot the step writes (nemo_hip_neuron_group.levels): a neuron's
		 * pass follows the passes of the lower-index sources it reads that
		 * slot from; indices ascend, so one pass in index order sets them */
		for(Group& g : m_groups) {
			const unsigned H = g.type.history;
			const bool reads = g.type.rcmSources || g.type.rcmDelays || g.type.rcmWeights || g.type.rcmForward;
			if(H < 2 || !reads || g.count == 0) continue;
			bool any = false;
			std::vector<uint32_t> level(g.count, 0);
			uint32_t levels = 1;
			for(unsigned i = 0; i < g.count; ++i) {
				const uint32_t tg = g.start + i;
				for(uint32_t k = ptr[tg]; k < ptr[tg + 1]; ++k) {
					const unsigned d = (sd[k] >> 26) + 1;
					if((d - 1) % H != H - 1) continue;
					any = true;
					const uint32_t src = (sd[k] & 0x3ffffffu) - g.start;
					if(src < i) level[i] = std::max(level[i], level[src] + 1);
				}
				levels = std::max(levels, level[i] + 1);
			}
			if(!any) continue;
			std::vector<uint32_t> lptr(levels + 1, 0), lidx(g.count);
			for(unsigned i = 0; i < g.count; ++i) lptr[level[i] + 1] += 1;
			for(unsigned l = 0; l < levels; ++l) lptr[l + 1] += lptr[l];
			std::vector<uint32_t> lfill(lptr.begin(), lptr.end() - 1);
			for(unsigned i = 0; i < g.count; ++i) lidx[lfill[level[i]]++] = i;
			g.levelPtr = lptr;
			g.levelIdx.allocate(lidx.size());
			HIP_CHECK(hipMemcpy(g.levelIdx.ptr, lidx.data(), g.levelIdx.bytes(), hipMemcpyHostToDevice));
			g.dev.levels = levels;
			g.dev.level_ptr = g.levelPtr.data();
			g.dev.level_idx = g.levelIdx.ptr;
			if(g.type.scratchPlanes < 1) {
				throw Error(NEMO_API_UNSUPPORTED, "neuron type " + g.type.name + " reads the state slot a step "
						"writes and needs a scratch plane for its old contents ([hip] scratch-planes)");
			}
		}
		m_rcmAllPtr.allocate(ptr.size());
		m_rcmAllSd.allocate(sd.size());
		m_rcmAllW.allocate(w.size());
		HIP_CHECK(hipMemcpy(m_rcmAllPtr.ptr, ptr.data(), m_rcmAllPtr.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(m_rcmAllSd.ptr, sd.data(), m_rcmAllSd.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(m_rcmAllW.ptr, w.data(), m_rcmAllW.bytes(), hipMemcpyHostToDevice));
		m_checkStatus = true;
	}
	if(m_writeOnly) {
		std::vector<uint32_t>().swap(m_synAddr);
	}
	m_rcmPtr.allocate(rcmPtr.size());
	HIP_CHECK(hipMemcpy(m_rcmPtr.ptr, rcmPtr.data(), m_rcmPtr.bytes(), hipMemcpyHostToDevice));
	if(m_stdp) {
		m_rcmSd.allocate(rcmSd.size());
		m_rcmFwd.allocate(rcmFwd.size());
		m_rcmAcc.allocate(rcmSd.size());
		HIP_CHECK(hipMemcpy(m_rcmSd.ptr, rcmSd.data(), m_rcmSd.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemcpy(m_rcmFwd.ptr, rcmFwd.data(), m_rcmFwd.bytes(), hipMemcpyHostToDevice));
		HIP_CHECK(hipMemset(m_rcmAcc.ptr, 0, m_rcmAcc.bytes()));
		m_rcmOf.allocate(m_fcm.n);
		HIP_CHECK(nemo_launch_invert_rcm(nullptr, m_rcmFwd.ptr, m_rcmOf.ptr, rcmFwd.size(), m_fcm.n));
		HIP_CHECK(hipDeviceSynchronize());
	}
}

/* StdpProcess constructor (StdpProcess.cpp:20-52) and StdpFunction::getBits
 * (StdpFunction.cpp:72-108): fixed-point tables and window bit masks. */
void Simulation::buildStdp(const Configuration& conf)
{
	std::memset(&m_propagate, 0, sizeof(m_propagate));
	std::memset(&m_stdpArgs, 0, sizeof(m_stdpArgs));
	if(!m_stdp) return;
	const unsigned fbits = m_fbits;
	m_minE = fxToFix(conf.minE, fbits);
	m_maxE = fxToFix(conf.maxE, fbits);
	m_minI = fxToFix(conf.minI, fbits);
	m_maxI = fxToFix(conf.maxI, fbits);
	nemo_hip_stdp& st = m_stdpArgs;
	for(size_t i = 0; i < conf.prefire.size(); ++i) st.fn_pre[i] = fxToFix(conf.prefire[i], fbits);
	for(size_t i = 0; i < conf.postfire.size(); ++i) st.fn_post[i] = fxToFix(conf.postfire[i], fbits);
	uint64_t pot = 0, dep = 0;
	unsigned n = 0;
	for(auto f = conf.postfire.rbegin(); f != conf.postfire.rend(); ++f, ++n) {
		if(*f > 0.0f) pot |= uint64_t(1) << n;
		if(*f < 0.0f) dep |= uint64_t(1) << n;
	}
	for(auto f = conf.prefire.begin(); f != conf.prefire.end(); ++f, ++n) {
		if(*f > 0.0f) pot |= uint64_t(1) << n;
		if(*f < 0.0f) dep |= uint64_t(1) << n;
	}
	const unsigned pre = unsigned(conf.prefire.size());
	const unsigned post = unsigned(conf.postfire.size());
	auto lowMask = [](unsigned k) { return k >= 64 ? ~uint64_t(0) : ((uint64_t(1) << k) - 1); };
	st.post_window = post;
	st.potentiation = pot;
	st.depression = dep;
	st.pre_post = post >= 64 ? 0 : (lowMask(pre) << post);
	st.post_pre = lowMask(post);
}

/* Split the (source tile, delay) pairs over B delivery blocks with about
 * equal synapse counts, and give each block its target-tile window.  The raw
 * window of a block (the targets of its source tiles) is widened so that the
 * lower and upper ends never decrease from one block to the next; then the
 * blocks whose windows hold a given tile form one contiguous range, which the
 * update kernels read. */
struct Partition {
	std::vector<uint32_t> bounds;  /* B+1 */
	std::vector<uint2> windows;    /* B */
	std::vector<uint32_t> range;   /* per tile: first | end << 16 */
	unsigned maxWindow = 0;
};

static Partition partitionPairs(const std::vector<uint64_t>& pairWeight, const std::vector<uint2>& srcWindow,
		unsigned D, unsigned tiles, unsigned dFirst, unsigned dCount, unsigned B)
{
	Partition P;
	if(B == 0 || dCount == 0) {
		P.range.assign(tiles, 0);
		return P;
	}
	/* pairs of the set: q = tile * dCount + d - dFirst */
	const size_t Q = size_t(tiles) * dCount;
	std::vector<uint64_t> weight(Q);
	for(size_t p = 0; p < tiles; ++p) {
		for(unsigned j = 0; j < dCount; ++j) weight[p * dCount + j] = pairWeight[p * D + dFirst - 1 + j];
	}
	uint64_t total = 0;
	for(uint64_t w : weight) total += w;
	P.bounds.assign(B + 1, 0);
	size_t q = 0;
	uint64_t cum = 0;
	for(unsigned b = 1; b < B; ++b) {
		const uint64_t target = (total * b + B - 1) / B;
		while(q < Q && cum < target) cum += weight[q++];
		P.bounds[b] = uint32_t(q);
	}
	P.bounds[B] = uint32_t(Q);

	std::vector<uint32_t> lo(B, tiles), hi(B, 0);
	for(unsigned b = 0; b < B; ++b) {
		if(P.bounds[b] == P.bounds[b + 1]) continue;
		const unsigned p0 = P.bounds[b] / dCount, p1 = (P.bounds[b + 1] - 1) / dCount;
		for(unsigned p = p0; p <= p1; ++p) {
			if(srcWindow[p].y == 0) continue;
			lo[b] = std::min(lo[b], srcWindow[p].x);
			hi[b] = std::max(hi[b], srcWindow[p].y);
		}
	}
	for(unsigned b = 1; b < B; ++b) hi[b] = std::max(hi[b], hi[b - 1]);
	for(unsigned b = B - 1; b-- > 0;) lo[b] = std::min(lo[b], lo[b + 1]);
	P.windows.resize(B);
	for(unsigned b = 0; b < B; ++b) {
		if(lo[b] > hi[b]) lo[b] = hi[b];
		P.windows[b] = make_uint2(lo[b], hi[b]);
		P.maxWindow = std::max(P.maxWindow, hi[b] - lo[b]);
	}
	P.range.assign(tiles, 0);
	unsigned first = 0, end = 0;
	for(unsigned t = 0; t < tiles; ++t) {
		while(first < B && hi[first] <= t) ++first;
		while(end < B && lo[end] <= t) ++end;
		P.range[t] = first < end ? (first | (end << 16)) : 0u;
	}
	return P;
}

/* Default shape: bin capacity 4096 events (halved while the LDS does not
 * allow it), enough delivery blocks for about 1% of the synapses per step,
 * and, with delays above 1 and few enough blocks to be resident together, a
 * delay-1 set of about 1/D of them, so that the update launches carry the
 * next step's delivery and push its delay-1 events (overlap 3); STDP
 * accumulated in batches beside the steps, one bitmap word per 64 threads
 * of an STDP block.  NEMO_HIP_DELIVER_BLOCKS,
 * NEMO_HIP_LATE_BLOCKS, NEMO_HIP_BIN_CAPACITY, NEMO_HIP_STDP_MODE,
 * NEMO_HIP_STDP_WORDS and NEMO_HIP_OVERLAP override it (tuning). */
LaunchShape Simulation::defaultShape() const
{
	LaunchShape s;
	const uint64_t pairs = uint64_t(m_D) * m_tiles;
	for(unsigned cap = 4096; cap >= 256; cap >>= 1) {
		double blocks = std::ceil(0.01 * double(m_nsynapses) / cap);
		blocks = std::max(2.0, std::min({blocks, double(NEMO_MAX_BIN_BLOCKS), double(pairs) + 1.0}));
		s.deliverBlocks = unsigned(blocks);
		s.lateBlocks = 0;
		s.binCapacity = cap;
		if(ldsNeeded(s) <= m_ldsBudget) break;
	}
	if(m_D >= 2) {
		/* one launch per step pays when the update launch's tiles and the
		 * delivery blocks it carries are all resident at once (a 100k-neuron
		 * network: 391 + 245 blocks, 3 per CU); with many more (a batch of
		 * networks, 4050 tiles) its blocks wait for each other's CU slots */
		LaunchShape c = s;
		c.lateBlocks = std::max(1u, unsigned(std::lround(double(s.deliverBlocks) / m_D)));
		c.deliverBlocks = std::min(s.deliverBlocks, NEMO_MAX_BIN_BLOCKS - c.lateBlocks);
		c.overlap = 3;
		const size_t lds = ldsNeeded(c);
		const size_t perCu = std::min<size_t>(m_ldsPerCu / std::max<size_t>(1, lds + COLAUNCH_LDS_MARGIN),
				2048 / NEMO_COLAUNCH_BLOCK);
		if(lds <= m_ldsBudget && size_t(m_tiles) + c.deliverBlocks <= m_cuCount * perCu) s = c;
	}
	s.stdpMode = 4;
	s.stdpWords = NEMO_STDP_BATCH_WORDS; /* about one batch block per CU disturbs the steps least */
	auto env = [](const char* name, int lo, int hi, unsigned& field) {
		if(const char* e = std::getenv(name)) field = unsigned(std::max(lo, std::min(std::atoi(e), hi)));
	};
	env("NEMO_HIP_DELIVER_BLOCKS", 0, NEMO_MAX_BIN_BLOCKS, s.deliverBlocks);
	env("NEMO_HIP_LATE_BLOCKS", 0, NEMO_MAX_BIN_BLOCKS, s.lateBlocks);
	env("NEMO_HIP_BIN_CAPACITY", 64, 32768, s.binCapacity);
	env("NEMO_HIP_STDP_MODE", 0, 4, s.stdpMode);
	env("NEMO_HIP_STDP_WORDS", 1, NEMO_STDP_BATCH_MAX_WORDS, s.stdpWords);
	if(s.stdpMode != 4) s.stdpWords = std::min<unsigned>(s.stdpWords, NEMO_STDP_MAX_WORDS);
	env("NEMO_HIP_OVERLAP", 0, 3, s.overlap);
	return s;
}

/* Dynamic LDS the delivery launches of a shape need (the larger set). */
size_t Simulation::ldsNeeded(const LaunchShape& s) const
{
	unsigned w = 1;
	const bool split = s.lateBlocks > 0;
	const unsigned dFirst = split ? 2 : 1, dCount = split ? m_D - 1 : m_D;
	if(dCount && s.deliverBlocks) {
		w = std::max(w, partitionPairs(m_pairWeight, m_srcWindow, m_D, m_tiles, dFirst, dCount, s.deliverBlocks).maxWindow);
	}
	if(split) w = std::max(w, partitionPairs(m_pairWeight, m_srcWindow, m_D, m_tiles, 1, 1, s.lateBlocks).maxWindow);
	size_t need = nemo_propagate_lds_bytes(w, s.binCapacity);
	if(split && s.overlap >= 2) {
		/* the update launches carry delivery blocks, which share their
		 * dynamic LDS with the update blocks */
		const size_t co = nemo_colaunch_lds_bytes(w, s.binCapacity, 1) + COLAUNCH_LDS_MARGIN;
		if(co > m_ldsBudget) need = std::max(need, co);
	}
	return need;
}

void Simulation::checkShape(const LaunchShape& s) const
{
	if(s.deliverBlocks + s.lateBlocks < 1 || s.deliverBlocks + s.lateBlocks > NEMO_MAX_BIN_BLOCKS) {
		throw Error(NEMO_INVALID_INPUT, "delivery blocks: 1 to " + std::to_string(NEMO_MAX_BIN_BLOCKS) + " in all");
	}
	if(s.deliverBlocks == 0 && (s.lateBlocks == 0 || m_D >= 2)) {
		throw Error(NEMO_INVALID_INPUT, "the delays delivered by the first block set need at least one block");
	}
	if(s.overlap > 3) {
		throw Error(NEMO_INVALID_INPUT, "overlap must be 0, 1, 2 or 3");
	}
	if(s.lateBlocks && s.overlap >= 2 && s.stdpMode == 2 && s.stdpWords > NEMO_COLAUNCH_BLOCK / 64) {
		throw Error(NEMO_INVALID_INPUT, "STDP blocks carried by the update launches cover at most "
				+ std::to_string(NEMO_COLAUNCH_BLOCK / 64) + " bitmap words");
	}
	if(s.binCapacity < 64 || s.binCapacity > 32768) {
		throw Error(NEMO_INVALID_INPUT, "bin capacity must be in [64, 32768]");
	}
	const unsigned maxWords = s.stdpMode == 4 ? NEMO_STDP_BATCH_MAX_WORDS : NEMO_STDP_MAX_WORDS;
	if(s.stdpWords < 1 || s.stdpWords > maxWords) {
		throw Error(NEMO_INVALID_INPUT, "STDP words per block must be in [1, " + std::to_string(maxWords)
				+ "] for this STDP mode");
	}
	if(s.stdpMode > 4) {
		throw Error(NEMO_INVALID_INPUT, "STDP mode must be 0, 1, 2, 3 or 4");
	}
}

/* Set up the buffers and kernel arguments of a launch shape.  Bins and table
 * only carry data from a step's delivery to the same step's update, so the
 * shape may change between any two steps. */
void Simulation::applyShape(const LaunchShape& s)
{
	checkShape(s);
	/* lateBlocks == 0: one set for all delays; else delays 2..D and delay 1 */
	const bool split = s.lateBlocks > 0;
	const unsigned dFirst = split ? 2 : 1, dCount = split ? m_D - 1 : m_D;
	const unsigned Be = dCount ? s.deliverBlocks : 0;
	const unsigned Bl = s.lateBlocks;
	const unsigned C = Be + Bl;
	const Partition Pe = partitionPairs(m_pairWeight, m_srcWindow, m_D, m_tiles, dFirst, dCount, Be);
	const Partition Pl = partitionPairs(m_pairWeight, m_srcWindow, m_D, m_tiles, 1, 1, Bl);
	const unsigned maxWindow = std::max({1u, Pe.maxWindow, Pl.maxWindow});
	if(nemo_propagate_lds_bytes(maxWindow, s.binCapacity) > m_ldsBudget
			|| (split && s.overlap >= 2
				&& nemo_colaunch_lds_bytes(maxWindow, s.binCapacity, 1) + COLAUNCH_LDS_MARGIN > m_ldsBudget)) {
		throw Error(NEMO_INVALID_INPUT, "launch shape needs more LDS than a block can have");
	}
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	clearGraphs();
	m_shape = s;
	m_shape.deliverBlocks = Be;

	/* both sets back to back: bounds [Be+1 | Bl+1], windows [Be | Bl]; the
	 * per-tile ranges in absolute table columns, early set then late set */
	std::vector<uint32_t> bounds(Pe.bounds);
	bounds.insert(bounds.end(), Pl.bounds.begin(), Pl.bounds.end());
	std::vector<uint2> windows(Pe.windows);
	windows.insert(windows.end(), Pl.windows.begin(), Pl.windows.end());
	std::vector<uint32_t> range(2 * size_t(m_tiles), 0);
	for(unsigned t = 0; t < m_tiles; ++t) {
		range[t] = Pe.range[t];
		const uint32_t r = Pl.range[t];
		if((r >> 16) > (r & 0xffffu)) range[m_tiles + t] = ((r & 0xffffu) + Be) | (((r >> 16) + Be) << 16);
	}
	m_pairBounds.allocate(std::max<size_t>(1, bounds.size()));
	HIP_CHECK(hipMemcpy(m_pairBounds.ptr, bounds.data(), bounds.size() * 4, hipMemcpyHostToDevice));
	m_windows.allocate(std::max<size_t>(1, windows.size()));
	HIP_CHECK(hipMemcpy(m_windows.ptr, windows.data(), windows.size() * sizeof(uint2), hipMemcpyHostToDevice));
	m_binRange.allocate(range.size());
	HIP_CHECK(hipMemcpy(m_binRange.ptr, range.data(), m_binRange.bytes(), hipMemcpyHostToDevice));
	/* the update reads only a tile's column ranges (one dependent load more)
	 * when they cover under a quarter of the table row on average;
	 * NEMO_HIP_BIN_RANGED=0 or 1 forces the choice */
	{
		uint64_t width = 0;
		for(uint32_t r : range) width += (r >> 16) - (r & 0xffffu);
		m_binRanged = width * 4 < uint64_t(m_tiles) * C;
		if(const char* e = std::getenv("NEMO_HIP_BIN_RANGED")) m_binRanged = std::atoi(e) != 0;
	}
	m_bins.allocate(2 * size_t(C) * s.binCapacity);
	m_binTable.allocate(2 * size_t(m_tiles) * C);
	HIP_CHECK(hipMemset(m_binTable.ptr, 0, m_binTable.bytes()));
	m_stdpBlocks = (m_stdp && m_nplastic) ? unsigned((m_firedWords + s.stdpWords - 1) / s.stdpWords) : 0;

	nemo_propagate_args& a = m_propagate;
	a.clock = m_clock.ptr;
	a.step_offset = 0;
	a.fired_bits = m_firedBits.ptr;
	a.fired_words = m_firedWords;
	a.deliver_blocks = 0;
	a.table_columns = C;
	a.max_window = maxWindow;
	a.rowptr = m_rowptr.ptr;
	a.fcm = m_fcm.ptr;
	a.D = m_D;
	a.tiles = m_tiles;
	a.bin_capacity = s.binCapacity;
	a.bin_table = m_binTable.ptr;
	a.bins = m_bins.ptr;
	a.acc = m_acc.ptr;
	a.acc_pitch = m_accPitch;
	a.block_events = m_blockEvents.ptr;
	a.stdp_words = m_stdpBlocks ? s.stdpWords : 0;
	a.hist = m_hist.ptr;
	a.hist_pitch = m_accPitch;
	a.stdp_snap = m_stdpSnap.ptr;
	a.stdp = m_stdpArgs;

	m_early = a;
	m_early.deliver_blocks = m_nsynapses ? Be : 0;
	m_early.block_base = 0;
	m_early.d_first = dFirst;
	m_early.d_count = std::max(1u, dCount);
	m_early.pair_bounds = m_pairBounds.ptr;
	m_early.windows = m_windows.ptr;
	m_late = a;
	m_late.deliver_blocks = m_nsynapses ? Bl : 0;
	m_late.block_base = Be;
	m_late.d_first = 1;
	m_late.d_count = 1;
	m_late.pair_bounds = m_pairBounds.ptr + Pe.bounds.size();
	m_late.windows = m_windows.ptr + Be;
	resetStdpBatches();
}

/* STDP batches start over from a snapshot of the history after the last
 * step taken, with nothing pending (every earlier step is accumulated). */
void Simulation::resetStdpBatches()
{
	if(!m_stdpSnap.ptr) return;
	if(m_hostStep == 0) {
		HIP_CHECK(hipMemsetAsync(m_stdpSnap.ptr, 0, m_stdpSnap.bytes(), m_stream));
	} else {
		const uint64_t* last = m_hist.ptr + NEMO_HIST_INDEX(0, m_hostStep - 1, m_accPitch);
		HIP_CHECK(hipMemcpy2DAsync(m_stdpSnap.ptr, 2 * sizeof(uint64_t), last, sizeof(uint64_t), sizeof(uint64_t),
					m_accPitch, hipMemcpyDeviceToDevice, m_stream));
	}
	copyScalar(&m_clock.ptr->stdp_snaps, 0);
	copyScalar(&m_clock.ptr->stdp_last, m_hostStep - 1);
	copyScalar(&m_clock.ptr->stdp_first, m_hostStep);
	copyScalar(&m_clock.ptr->stdp_count, 0);
	HIP_CHECK(hipStreamSynchronize(m_stream));
}

void Simulation::setLaunchShape(const LaunchShape& s)
{
	flushStdp();
	applyShape(s);
}

void Simulation::allocateStepState()
{
	const size_t N = m_npad;
	int maxLds = 65536;
	(void) hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, m_device);
	m_ldsBudget = size_t(maxLds) - std::min<size_t>(size_t(maxLds), nemo_propagate_static_lds());
	int cus = 256, ldsPerCu = 163840;
	(void) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m_device);
	(void) hipDeviceGetAttribute(&ldsPerCu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, m_device);
	m_cuCount = unsigned(std::max(1, cus));
	m_ldsPerCu = size_t(std::max(65536, ldsPerCu));
	m_clock.allocate(1);
	nemo_hip_clock c0;
	c0.base_step = 0;
	c0.cycle_origin = 0;
	c0.stdp_flushed = ~uint64_t(0);
	HIP_CHECK(hipMemcpy(m_clock.ptr, &c0, sizeof(c0), hipMemcpyHostToDevice));
	m_firedWords = N / 64;
	m_firedBits.allocate(size_t(NEMO_STEP_RING) * m_firedWords);
	HIP_CHECK(hipMemset(m_firedBits.ptr, 0, m_firedBits.bytes()));
	m_accPitch = N;
	m_acc.allocate(4 * m_accPitch); /* [step parity][E, I][N] */
	HIP_CHECK(hipMemset(m_acc.ptr, 0, m_acc.bytes()));
	m_hist.allocate(2 * N); /* plane s % 2 holds the history after step s */
	if(m_stdp && m_nplastic) {
		/* two history snapshots per neuron, from which an STDP batch
		 * accumulates up to 64 steps between them */
		m_stdpSnap.allocate(2 * N);
	}
	HIP_CHECK(hipMemset(m_hist.ptr, 0, m_hist.bytes()));
	m_ext.allocate(N);
	HIP_CHECK(hipMemset(m_ext.ptr, 0, m_ext.bytes()));
	m_fstim.allocate((N + 31) / 32);
	HIP_CHECK(hipMemset(m_fstim.ptr, 0, m_fstim.bytes()));
	m_status.allocate(1);
	HIP_CHECK(hipMemset(m_status.ptr, 0, m_status.bytes()));
	m_debugCurrent.allocate(2 * m_accPitch);
	HIP_CHECK(hipMemset(m_debugCurrent.ptr, 0, m_debugCurrent.bytes()));
	HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_pinnedBits),
				size_t(NEMO_STEP_RING) * m_firedWords * sizeof(uint64_t), hipHostMallocDefault));
	/* touched now, so that the first readout of a page does not fault it in */
	std::memset(m_pinnedBits, 0, size_t(NEMO_STEP_RING) * m_firedWords * sizeof(uint64_t));

	/* work counters, sized for any launch shape */
	/* per table column, then per 64 device indices (events pushed by updates) */
	m_blockEvents.allocate(NEMO_MAX_BIN_BLOCKS + m_npad / 64);
	HIP_CHECK(hipMemset(m_blockEvents.ptr, 0, m_blockEvents.bytes()));
	HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_pinnedCounters), 2 * m_blockEvents.bytes(),
				hipHostMallocDefault));
	std::memset(m_pinnedCounters, 0, 2 * m_blockEvents.bytes());
	m_blockExamined.allocate(std::max<size_t>(1, m_firedWords));
	HIP_CHECK(hipMemset(m_blockExamined.ptr, 0, m_blockExamined.bytes()));

	if(m_stdp && m_nplastic) {
		m_stdpArgs.rcm_ptr = m_rcmPtr.ptr;
		m_stdpArgs.rcm_sd = m_rcmSd.ptr;
		m_stdpArgs.rcm_acc = m_rcmAcc.ptr;
		m_stdpArgs.examined = m_blockExamined.ptr;
	} else {
		m_stdpArgs.rcm_ptr = nullptr;
	}
}

nemo_hip_step_buffers Simulation::stepBuffers(uint64_t offset, bool ext, bool fstim) const
{
	nemo_hip_step_buffers b;
	std::memset(&b, 0, sizeof(b));
	b.step_offset = offset;
	b.clock = m_clock.ptr;
	b.bin_table = m_binTable.ptr;
	b.bin_range = m_binRange.ptr;
	b.bins = m_bins.ptr;
	b.bin_blocks = m_nsynapses ? m_shape.deliverBlocks + m_shape.lateBlocks : 0;
	b.bin_capacity = m_shape.binCapacity;
	b.tiles = m_tiles;
	b.bin_ranged = m_binRanged ? 1u : 0u;
	b.acc = m_acc.ptr;
	b.acc_pitch = m_accPitch;
	b.ext_current = ext ? m_ext.ptr : nullptr;
	b.fstim = fstim ? m_fstim.ptr : nullptr;
	b.hist = m_hist.ptr;
	b.hist_pitch = m_accPitch;
	b.fired_bits = m_firedBits.ptr;
	b.fired_words = m_firedWords;
	b.fbits = m_fbits;
	b.debug_current = m_debug ? m_debugCurrent.ptr : nullptr;
	b.stdp = m_stdpArgs;
	if(m_shape.stdpMode != 1) b.stdp.rcm_ptr = nullptr; /* STDP runs in a launch of its own */
	b.rcm_all.ptr = m_rcmAllPtr.ptr;
	b.rcm_all.source_delay = m_rcmAllSd.ptr;
	b.rcm_all.weight = m_rcmAllW.ptr;
	b.status = m_status.ptr;
	b.ext_held = m_ext.ptr;
	return b;
}

float* Simulation::statePlane(const Group& g, unsigned var) const
{
	const unsigned H = g.type.history;
	const size_t slot = H > 1 && m_updated ? size_t((m_lastCycle + 1) % H) : 0;
	return g.state.ptr + (slot * g.type.nState + var) * g.count;
}

/* Report what the kernels flagged (NEMO_HIP_STATUS_*) since the last check;
 * the stream must be synchronised. */
void Simulation::checkStatus()
{
	if(!m_checkStatus) return;
	unsigned st = 0;
	HIP_CHECK(hipMemcpy(&st, m_status.ptr, sizeof(st), hipMemcpyDeviceToHost));
	if(!st) return;
	HIP_CHECK(hipMemset(m_status.ptr, 0, sizeof(unsigned)));
	if(st & NEMO_HIP_STATUS_SINF_RANGE) {
		throw Error(NEMO_API_UNSUPPORTED, "a neuron model took the sine of an argument outside (-120, 120), "
				"which the HIP backend does not evaluate as the C library does; the step's state is not the reference's");
	}
}

/* Copy a host scalar to the device on the stream.  The source is a fresh
 * pinned slot of a ring; before the ring wraps, the stream is synchronised,
 * so every copy queued from a slot has run before the slot is written again
 * (copy j reuses the slot of copy j - SCALAR_SLOTS, which precedes that
 * synchronisation).  Never called while the stream is being captured. */
void Simulation::copyScalar(uint64_t* dst, uint64_t value)
{
	if(m_scalarNext % SCALAR_SLOTS == 0 && m_scalarNext) {
		HIP_CHECK(hipStreamSynchronize(m_stream));
	}
	uint64_t* slot = &m_pinnedScalars[m_scalarNext++ % SCALAR_SLOTS];
	*slot = value;
	HIP_CHECK(hipMemcpyAsync(dst, slot, sizeof(uint64_t), hipMemcpyHostToDevice, m_stream));
}

/* Wait for an event (or a stream) by polling it for up to 50 ms before the
 * blocking wait: a blocking wait can put the thread to sleep and wake it
 * milliseconds after the work ended, which costs more than a whole short
 * run of steps. */
static void pollThenWait(hipError_t (*query)(void*), hipError_t (*wait)(void*), void* x)
{
	const auto t0 = std::chrono::steady_clock::now();
	for(;;) {
		const hipError_t e = query(x);
		if(e == hipSuccess) return;
		if(e != hipErrorNotReady) HIP_CHECK(e);
		if(std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
#if defined(__x86_64__)
		__builtin_ia32_pause();
#endif
	}
	HIP_CHECK(wait(x));
}

void Simulation::waitEvent(hipEvent_t e)
{
	pollThenWait([](void* x) { return hipEventQuery(static_cast<hipEvent_t>(x)); },
			[](void* x) { return hipEventSynchronize(static_cast<hipEvent_t>(x)); }, e);
}

void Simulation::waitStream(hipStream_t s)
{
	pollThenWait([](void* x) { return hipStreamQuery(static_cast<hipStream_t>(x)); },
			[](void* x) { return hipStreamSynchronize(static_cast<hipStream_t>(x)); }, s);
}

/* The step base used by the next launches: step = base + launch offset. */
void Simulation::setBase(uint64_t base)
{
	copyScalar(&m_clock.ptr->base_step, base);
	m_base = base;
}

void Simulation::launchDelivery(const nemo_propagate_args& set, uint64_t offset, hipStream_t st, bool withStdp,
		hipEvent_t start, hipEvent_t stop)
{
	const unsigned stdpBlocks = withStdp ? m_stdpBlocks : 0;
	if(!set.deliver_blocks && !stdpBlocks) {
		if(start) HIP_CHECK(hipEventRecord(start, st));
		if(stop) HIP_CHECK(hipEventRecord(stop, st));
		return;
	}
	nemo_propagate_args a = set;
	a.step_offset = offset;
	HIP_CHECK(nemo_launch_propagate(st, &a, stdpBlocks, 1, start, stop));
}

/* Steps first..first+k-1 on the stream (SimulationBackend.cpp:22-77): spike
 * delivery (delays 2..D, then delay 1), the neuron updates, and the STDP
 * accumulation of the previous step.  Within a run of steps, the delivery of
 * delays 2..D for step t+2 is forked onto a branch right after the update of
 * step t (it needs firing up to step t and the buffers of its parity, which
 * the update of step t has finished with) and joined before the update of
 * step t+2, so it runs beside the update of step t+1 and the delay-1
 * delivery of steps t+1 and t+2; the first step of a run delivers its delays
 * 2..D in order.  STDP accumulation of step t-1 runs inside the update
 * kernels, shares the delivery launch of step t, or runs on its own branch
 * forked after the update of step t-1 and joined either before the update of
 * step t or, with one step of slack, before the update of step t+1: it reads
 * the history plane of step t-1, which the update of step t+1 rewrites, and
 * the accumulators it updates are serialised on the branch's stream.  The
 * branch of the last step is joined at the end of the run.  With overlap 2
 * the update launch of step t carries the delivery of delays 2..D of step
 * t+1 instead, and with overlap 3 its tiles also push their delay-1 events
 * for step t+1, whose update then reads only the first set's bins.  Within
 * the run only: the next run (after an applyStdp, say) delivers its first
 * step from scratch, with the weights of that time.  In STDP mode 4
 * the pending batch runs on a branch beside all the steps and a history
 * snapshot ends the run (and every 64 steps).  With ev (four events per
 * step), the kernels record the start and end of the step's delivery on this
 * stream and of its neuron updates. */
void Simulation::enqueueSteps(uint64_t first, unsigned k, bool ext, bool fstim, hipEvent_t* ev, bool forkPending,
		bool advanceBase)
{
	const bool split = m_late.deliver_blocks && m_early.deliver_blocks;
	const bool overlap = split && m_shape.overlap == 1 && k > 1;
	/* the update of step t carries the delivery of the delays 2..D of step
	 * t+1 as extra blocks (and, with fused STDP, the STDP blocks); with
	 * overlap 3 its tiles also push the delay-1 events of their firing into
	 * the accumulators of step t+1, which then needs no delivery launch */
	const bool colaunch = split && m_shape.overlap >= 2;
	const bool pushed = split && m_shape.overlap == 3;
	/* STDP shares the delivery launch that follows the previous update, or
	 * the update launch when that carries propagate blocks */
	const bool fused = m_shape.stdpMode == 2 && m_stdpBlocks;
	const bool lagged = m_shape.stdpMode == 3 && m_stdpBlocks;
	const bool batched = m_shape.stdpMode == 4 && m_stdpBlocks;
	/* the pending STDP batch runs on its own branch beside the steps; it
	 * reads only the history snapshots, and is joined before the next
	 * snapshot replaces the older of them */
	bool forked = false;
	auto forkBatch = [&]() {
		HIP_CHECK(hipEventRecord(m_stdpFork, m_stream));
		HIP_CHECK(hipStreamWaitEvent(m_stdpStream, m_stdpFork, 0));
		HIP_CHECK(nemo_launch_stdp_batch(m_stdpStream, &m_propagate, m_stdpBlocks));
		HIP_CHECK(hipEventRecord(m_stdpDone, m_stdpStream));
		forked = true;
	};
	/* without forkPending the caller knows that no batch is pending */
	if(batched && k && forkPending) forkBatch();
	bool pending[2] = {false, false};
	auto forkEarly = [&](unsigned i) {
		HIP_CHECK(hipEventRecord(m_earlyFork, m_stream));
		HIP_CHECK(hipStreamWaitEvent(m_earlyStream, m_earlyFork, 0));
		launchDelivery(m_early, first + i, m_earlyStream, false);
		HIP_CHECK(hipEventRecord(m_earlyDone[i & 1], m_earlyStream));
		pending[i & 1] = true;
	};
	for(unsigned i = 0; i < k; ++i) {
		const uint64_t off = first + i;
		if(m_shape.stdpMode == 0 && m_stdpBlocks) {
			joinStdp();
			HIP_CHECK(hipEventRecord(m_stdpFork, m_stream));
			HIP_CHECK(hipStreamWaitEvent(m_stdpStream, m_stdpFork, 0));
			nemo_propagate_args sa = m_propagate;
			sa.step_offset = off;
			HIP_CHECK(nemo_launch_propagate(m_stdpStream, &sa, m_stdpBlocks, 0, nullptr, nullptr));
			HIP_CHECK(hipEventRecord(m_stdpDone, m_stdpStream));
			m_stdpPending = true;
		}
		if(lagged) {
			HIP_CHECK(hipEventRecord(m_stdpFork, m_stream));
			HIP_CHECK(hipStreamWaitEvent(m_stdpStream, m_stdpFork, 0));
			nemo_propagate_args sa = m_propagate;
			sa.step_offset = off;
			HIP_CHECK(nemo_launch_propagate(m_stdpStream, &sa, m_stdpBlocks, 0, nullptr, nullptr));
			HIP_CHECK(hipEventRecord(m_stdpLag[i & 1], m_stdpStream));
		}
		/* with ev: four events per step, recorded by the kernels: start
		 * and end of the delivery on this stream, of the neuron updates */
		hipEvent_t ps = ev ? ev[4 * i] : nullptr, pe = ev ? ev[4 * i + 1] : nullptr;
		if(split && pushed && i > 0) {
			/* delivered by the previous step's update launch (no propagate
			 * launch to time: ps and pe stay unrecorded) */
		} else if(split) {
			const bool early = i == 0 || (!overlap && !colaunch);
			if(early && !(i == 0 && overlap)) {
				/* both sets in one launch */
				nemo_propagate_args a = m_early, a2 = m_late;
				a.step_offset = a2.step_offset = off;
				HIP_CHECK(nemo_launch_propagate_sets(m_stream, &a, &a2, fused && !colaunch ? m_stdpBlocks : 0, ps,
							pe));
			} else {
				if(early) launchDelivery(m_early, off, m_stream, false, ps, nullptr);
				if(i == 0 && overlap) forkEarly(1);
				launchDelivery(m_late, off, m_stream, fused && !colaunch, early ? nullptr : ps, pe);
			}
		} else {
			launchDelivery(m_early, off, m_stream, fused, ps, pe);
		}
		if(pending[i & 1] && i > 0) {
			HIP_CHECK(hipStreamWaitEvent(m_stream, m_earlyDone[i & 1], 0));
			pending[i & 1] = false;
		}
		/* the STDP branch overlaps delivery only: beside the update it slows
		 * the update, which is on the step's critical path */
		joinStdp();
		if(lagged && i > 0) HIP_CHECK(hipStreamWaitEvent(m_stream, m_stdpLag[(i - 1) & 1], 0));
		nemo_hip_step_buffers b = stepBuffers(off, ext, fstim);
		if(colaunch) {
			b.co.args = m_early;
			b.co.args.step_offset = off + 1;
			b.co.deliver_blocks = i + 1 < k ? m_early.deliver_blocks : 0;
			b.co.stdp_blocks = fused ? m_stdpBlocks : 0;
			b.co.lds_bytes = b.co.deliver_blocks || b.co.stdp_blocks
				? unsigned(nemo_colaunch_lds_bytes(m_early.max_window, m_early.bin_capacity, b.co.deliver_blocks > 0))
				: 0;
			b.co.push = pushed && i + 1 < k ? 1u : 0u;
			b.co.push_rows = m_d1Rows.ptr;
		}
		if(pushed && i > 0) b.bin_first_set = m_early.deliver_blocks;
		size_t last = m_groups.size();
		for(size_t gi = 0; gi < m_groups.size(); ++gi) {
			if(m_groups[gi].dev.count) last = gi;
		}
		if(m_groupStreams.empty()) {
			if(ev) b.time_start = ev[4 * i + 2];
			for(size_t gi = 0; gi < m_groups.size(); ++gi) {
				Group& g = m_groups[gi];
				if(g.dev.count == 0) continue;
				if(ev && gi == last) b.time_stop = ev[4 * i + 3];
				HIP_CHECK(g.plugin->update(m_stream, &g.dev, &b));
				b.co.deliver_blocks = 0; /* the first group's launch carries them */
				b.co.stdp_blocks = 0;
				b.co.lds_bytes = 0;
				b.time_start = nullptr;
			}
			if(ev && last == m_groups.size()) {
				HIP_CHECK(hipEventRecord(ev[4 * i + 2], m_stream));
				HIP_CHECK(hipEventRecord(ev[4 * i + 3], m_stream));
			}
		} else {
			/* the models' updates touch disjoint tiles: the first on this
			 * stream (with any propagate blocks the launch carries), the
			 * others on branches joined before the step ends */
			if(ev) HIP_CHECK(hipEventRecord(ev[4 * i + 2], m_stream));
			HIP_CHECK(hipEventRecord(m_groupFork, m_stream));
			size_t side = 0;
			bool first = true;
			for(size_t gi = 0; gi < m_groups.size(); ++gi) {
				Group& g = m_groups[gi];
				if(g.dev.count == 0) continue;
				if(first) {
					HIP_CHECK(g.plugin->update(m_stream, &g.dev, &b));
					b.co.deliver_blocks = 0; /* the first group's launch carries them */
					b.co.stdp_blocks = 0;
					b.co.lds_bytes = 0;
					first = false;
					continue;
				}
				hipStream_t q = m_groupStreams[side];
				HIP_CHECK(hipStreamWaitEvent(q, m_groupFork, 0));
				HIP_CHECK(g.plugin->update(q, &g.dev, &b));
				HIP_CHECK(hipEventRecord(m_groupDone[side], q));
				++side;
			}
			for(size_t q = 0; q < side; ++q) HIP_CHECK(hipStreamWaitEvent(m_stream, m_groupDone[q], 0));
			if(ev) HIP_CHECK(hipEventRecord(ev[4 * i + 3], m_stream));
		}
		if(overlap && i + 2 < k) forkEarly(i + 2);
		if(batched && (i + 1 == k || (i + 1) % 64 == 0)) {
			/* snapshot after this step: the steps since the last one are
			 * pending for the next batch */
			if(forked) HIP_CHECK(hipStreamWaitEvent(m_stream, m_stdpDone, 0));
			/* the last step's snapshot is the last work of the steps: with
			 * advanceBase it moves the step base past them */
			const uint64_t advance = advanceBase && i + 1 == k ? first + k : 0;
			HIP_CHECK(nemo_launch_stdp_snapshot(m_stream, &m_propagate, first + i, advance));
			if(i + 1 < k) forkBatch();
		}
	}
	if(lagged && k) HIP_CHECK(hipStreamWaitEvent(m_stream, m_stdpLag[(k - 1) & 1], 0));
}

/* k steps with neuron updates were enqueued after m_hostStep */
void Simulation::advanceSteps(uint64_t k)
{
	if(!k) return;
	m_hostStep += k;
	m_lastCycle = m_hostStep - 1 - m_cycleOrigin;
	m_updated = true;
}

/* Make the main stream wait for the STDP branch of the last enqueued step. */
void Simulation::joinStdp()
{
	if(!m_stdpPending) return;
	HIP_CHECK(hipStreamWaitEvent(m_stream, m_stdpDone, 0));
	m_stdpPending = false;
}

/* Run the pending STDP accumulation of the last step on its own, and tell the
 * next propagate launch to skip it. */
void Simulation::flushStdp()
{
	if(enqueueFlush()) HIP_CHECK(hipStreamSynchronize(m_stream));
}

/* flushStdp's work queued on the stream without waiting for it; false if
 * nothing was pending. */
bool Simulation::enqueueFlush()
{
	joinStdp();
	if(!m_stdpBlocks || m_hostStep == 0 || m_flushed == m_hostStep - 1) return false;
	if(m_shape.stdpMode == 4) {
		/* the pending batch: every run of steps ends with a snapshot.  Here
		 * nothing runs beside it, so it takes its own kernel (512-thread
		 * blocks, 8 loads in flight per thread, 4 words per block); a kernel
		 * after it marks it done (the flushed step relative to the device's
		 * base, modulo 2^64) */
		nemo_propagate_args a = m_propagate;
		a.stdp_words = std::min(a.stdp_words, 4u);
		const unsigned blocks = unsigned((m_firedWords + a.stdp_words - 1) / a.stdp_words);
		HIP_CHECK(nemo_launch_stdp_flush(m_stream, &a, blocks, m_hostStep - 1 - m_base));
		m_flushed = m_hostStep - 1;
		return true;
	}
	const uint64_t s = m_hostStep - 1;
	nemo_propagate_args a = m_propagate;
	a.deliver_blocks = 0;
	a.step_offset = 0;
	setBase(s + 1);
	HIP_CHECK(nemo_launch_propagate(m_stream, &a, m_stdpBlocks, 1, nullptr, nullptr));
	copyScalar(&m_clock.ptr->stdp_flushed, s);
	m_flushed = s;
	return true;
}

/* Neurons (user indices) set in the firing bitmap of one step, in simulation
 * order (cpu/Simulation.cpp:200-209). */
static inline unsigned lowestBit(uint64_t w)
{
	return w ? unsigned(__builtin_ctzll(w)) : 0u;
}

template<typename Out>
static inline void appendFired(Out& out, const unsigned* p, size_t n)
{
	out.insert(out.end(), p, p + n);
}

template<>
inline void appendFired(Simulation::FiredSink& out, const unsigned* p, size_t n)
{
	out.append(p, n);
}

template<typename Out>
void Simulation::decodeFiring(const uint64_t* bits, Out& out) const
{
	/* into a buffer with four entries of slack, then appended in one go.  In
	 * a word whose neurons have consecutive user indices, its first four set
	 * bits are written unconditionally (the slots past its count are
	 * overwritten by the next word): most words hold at most four, so the
	 * loop's branches follow the words, not their random bits */
	m_decodeBuf.resize(m_firedWords * 64 + 4);
	unsigned* d = m_decodeBuf.data();
	size_t pos = 0;
	for(size_t w = 0; w < m_firedWords; ++w) {
		uint64_t word = bits[w];
		const unsigned base = m_wordBase[w];
		if(base != ~0u) {
			const unsigned cnt = unsigned(__builtin_popcountll(word));
			unsigned* o = d + pos;
			o[0] = base + lowestBit(word);
			word &= word - 1;
			o[1] = base + lowestBit(word);
			word &= word - 1;
			o[2] = base + lowestBit(word);
			word &= word - 1;
			o[3] = base + lowestBit(word);
			word &= word - 1;
			for(unsigned i = 4; word; ++i) {
				o[i] = base + unsigned(__builtin_ctzll(word));
				word &= word - 1;
			}
			pos += cnt;
		} else {
			while(word) {
				const unsigned bit = unsigned(__builtin_ctzll(word));
				word &= word - 1;
				d[pos++] = m_l2g[w * 64 + bit];
			}
		}
	}
	appendFired(out, d, pos);
}

/* Fired neurons of every network of a batch for one step: out gets each
 * network's user indices in turn, in that network's simulation order, and
 * counts one entry per network. */
template<typename Out>
void Simulation::decodeFiringBatch(const uint64_t* bits, Out& out, std::vector<unsigned>& counts) const
{
	if(m_nets == 1) {
		const size_t before = out.size();
		decodeFiring(bits, out);
		counts.push_back(unsigned(out.size() - before));
		return;
	}
	if(m_decodeInOrder) {
		const size_t c0 = counts.size();
		counts.resize(c0 + m_nets, 0u);
		for(size_t w = 0; w < m_firedWords; ++w) {
			uint64_t word = bits[w];
			while(word) {
				const unsigned l = unsigned(w * 64 + unsigned(__builtin_ctzll(word)));
				word &= word - 1;
				out.push_back(m_l2g[l]);
				++counts[c0 + m_netOf[l]];
			}
		}
		return;
	}
	std::vector<std::vector<unsigned> >& per = m_decodePer; /* kept between steps: no allocation per step */
	per.resize(m_nets);
	for(std::vector<unsigned>& v : per) v.clear();
	for(size_t w = 0; w < m_firedWords; ++w) {
		uint64_t word = bits[w];
		while(word) {
			const unsigned bit = unsigned(__builtin_ctzll(word));
			word &= word - 1;
			per[m_netOf[w * 64 + bit]].push_back(unsigned(w * 64 + bit));
		}
	}
	for(unsigned k = 0; k < m_nets; ++k) {
		std::vector<unsigned>& v = per[k];
		if(m_reorder[k]) {
			std::sort(v.begin(), v.end(), [this](unsigned x, unsigned y) { return m_simOf[x] < m_simOf[y]; });
		}
		for(unsigned l : v) out.push_back(m_l2g[l]);
		counts.push_back(unsigned(v.size()));
	}
}

/* Copy the bitmaps of steps [t0, t0+k) (k <= NEMO_STEP_RING) to m_pinnedBits
 * (or dst), on the simulation's stream (or st). */
void Simulation::readBitmaps(uint64_t t0, unsigned k, uint64_t* dst, hipStream_t st)
{
	const size_t W = m_firedWords;
	if(!dst) dst = m_pinnedBits;
	if(!st) st = m_stream;
	unsigned done = 0;
	while(done < k) {
		const uint64_t slot = (t0 + done) & (NEMO_STEP_RING - 1);
		const unsigned len = unsigned(std::min<uint64_t>(k - done, NEMO_STEP_RING - slot));
		HIP_CHECK(hipMemcpyAsync(dst + size_t(done) * W, m_firedBits.ptr + slot * W,
					size_t(len) * W * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
		done += len;
	}
}

const std::vector<unsigned>& Simulation::step(const unsigned* fstim, size_t nf,
		const unsigned* istimIdx, const float* istimVal, size_t ni)
{
	activate();
	if(m_nets > 1) {
		throw Error(NEMO_API_UNSUPPORTED, "a batch of networks is advanced with nemo_hip_run");
	}
	/* current stimulus, then firing stimulus (SimulationBackend.cpp:61-77);
	 * for repeated neurons the last current wins (cpu/Simulation.cpp:140-145) */
	std::vector<uint32_t> iIdx;
	std::vector<float> iVal;
	if(ni) {
		std::unordered_map<unsigned, size_t> last;
		for(size_t i = 0; i < ni; ++i) last[localIdx(istimIdx[i])] = i;
		for(auto& kv : last) {
			iIdx.push_back(kv.first);
			iVal.push_back(istimVal[kv.second]);
		}
	}
	std::vector<uint32_t> fIdx(nf);
	for(size_t i = 0; i < nf; ++i) fIdx[i] = localIdx(fstim[i]);

	const size_t nstim = std::max(iIdx.size(), fIdx.size());
	if(nstim > m_stimIdx.n) {
		HIP_CHECK(hipStreamSynchronize(m_stream));
		m_stimIdx.allocate(nstim);
		m_stimVal.allocate(nstim);
	}
	if(!iIdx.empty()) {
		HIP_CHECK(hipMemcpyAsync(m_stimIdx.ptr, iIdx.data(), iIdx.size() * 4, hipMemcpyHostToDevice, m_stream));
		HIP_CHECK(hipMemcpyAsync(m_stimVal.ptr, iVal.data(), iVal.size() * 4, hipMemcpyHostToDevice, m_stream));
		HIP_CHECK(nemo_launch_istim(m_stream, m_ext.ptr, m_stimIdx.ptr, m_stimVal.ptr, iIdx.size()));
		/* the index buffer is reused for the firing stimulus below */
		HIP_CHECK(hipStreamSynchronize(m_stream));
	}
	if(!fIdx.empty()) {
		HIP_CHECK(hipMemcpyAsync(m_stimIdx.ptr, fIdx.data(), fIdx.size() * 4, hipMemcpyHostToDevice, m_stream));
		HIP_CHECK(nemo_launch_fstim(m_stream, m_fstim.ptr, m_stimIdx.ptr, fIdx.size(), 1));
	}
	const uint64_t t = m_hostStep;
	setBase(t);
	enqueueSteps(0, 1, !iIdx.empty(), !fIdx.empty());
	advanceSteps(1);
	if(!fIdx.empty()) {
		HIP_CHECK(nemo_launch_fstim(m_stream, m_fstim.ptr, m_stimIdx.ptr, fIdx.size(), 0));
	}
	readBitmaps(t, 1);
	waitStream(m_stream);
	checkStatus();
	m_firedOut.clear();
	decodeFiring(m_pinnedBits, m_firedOut);
	return m_firedOut;
}

/* With batched STDP every graph of steps ends with a snapshot, which then
 * moves the step base on (the other placements set it from the host). */
bool Simulation::graphsAdvanceBase() const
{
	return m_shape.stdpMode == 4 && m_stdpBlocks;
}

hipGraphExec_t Simulation::graphFor(unsigned k, bool forkPending)
{
	const unsigned key = 2 * k + (forkPending ? 1u : 0u);
	auto it = m_graphs.find(key);
	if(it != m_graphs.end()) return it->second;
	hipGraph_t graph;
	HIP_CHECK(hipStreamBeginCapture(m_stream, hipStreamCaptureModeThreadLocal));
	const bool debug = m_debug;
	m_debug = false;
	try {
		enqueueSteps(0, k, false, false, nullptr, forkPending, graphsAdvanceBase());
	} catch(...) {
		m_debug = debug;
		m_stdpPending = false;
		(void) hipStreamEndCapture(m_stream, &graph);
		throw;
	}
	m_debug = debug;
	HIP_CHECK(hipStreamEndCapture(m_stream, &graph));
	hipGraphExec_t exec;
	HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
	HIP_CHECK(hipGraphDestroy(graph));
	/* its launch packets go to the device now, not in its first (timed) launch */
	HIP_CHECK(hipGraphUpload(exec, m_stream));
	m_graphs[key] = exec;
	return exec;
}

void Simulation::clearGraphs()
{
	for(auto& kv : m_graphs) (void) hipGraphExecDestroy(kv.second);
	m_graphs.clear();
}

uint64_t Simulation::sumCounters(const DeviceArray<unsigned long long>& a)
{
	std::vector<unsigned long long> h(a.n);
	HIP_CHECK(hipMemcpy(h.data(), a.ptr, a.bytes(), hipMemcpyDeviceToHost));
	uint64_t s = 0;
	for(unsigned long long v : h) s += v;
	return s;
}

/* The device-compacted firing readout of run() for a single network, set
 * up on first use; false (bitmaps decoded on the host instead) for batches
 * or when its lists would take more than 1 GiB. */
bool Simulation::ensureListReadout()
{
	if(m_nets != 1) return false;
	const size_t CHUNK = NEMO_STEP_RING / 2;
	if(m_listDev.n) return true;
	if(2 * CHUNK * m_npad * sizeof(uint32_t) > (size_t(1) << 30)) return false;
	m_l2gDev.allocate(m_npad);
	HIP_CHECK(hipMemcpy(m_l2gDev.ptr, m_l2g.data(), m_l2gDev.bytes(), hipMemcpyHostToDevice));
	m_listDev.allocate(2 * CHUNK * m_npad);
	m_listCountsDev.allocate(2 * CHUNK);
	HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m_pinnedListCounts), 2 * CHUNK * sizeof(uint32_t),
				hipHostMallocDefault));
	std::memset(m_pinnedListCounts, 0, 2 * CHUNK * sizeof(uint32_t));
	for(int h = 0; h < 2; ++h) growPinnedList(h, std::max<size_t>(65536, CHUNK * m_npad / 32), 0);
	return true;
}

/* Room for need entries in half h's pinned list, keeping its first keep. */
void Simulation::growPinnedList(int h, size_t need, size_t keep)
{
	if(need <= m_pinnedListCap[h]) return;
	const size_t cap = std::max(need, 2 * m_pinnedListCap[h]);
	uint32_t* p = nullptr;
	HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), cap * sizeof(uint32_t), hipHostMallocDefault));
	std::memset(p, 0, cap * sizeof(uint32_t)); /* its pages faulted in now, not by the readout */
	if(keep) std::memcpy(p, m_pinnedList[h], keep * sizeof(uint32_t));
	if(m_pinnedList[h]) HIP_CHECK(hipHostFree(m_pinnedList[h]));
	m_pinnedList[h] = p;
	m_pinnedListCap[h] = cap;
}

/* Steps without stimulus, captured in graphs of up to 64 steps; the firing
 * of each graph's steps is read back after it: for a single network as
 * per-step counts and a fired list compacted on the device (the list's
 * expected length copied with the counts, any rest once they are known),
 * for a batch as bitmaps decoded on the host. */
void Simulation::run(unsigned nsteps, unsigned stdpPeriod, float reward, bool readRaster,
		std::vector<unsigned>* counts, FiredSink* fired, uint64_t* events)
{
	activate();
	if(stdpPeriod && !m_stdp) {
		throw Error(NEMO_LOGIC_ERROR, "applyStdp called, but no STDP model specified");
	}
	waitStream(m_stream);
	/* the event counters before and after the run's work, copied on its
	 * stream (no blocking copy at either end) */
	/* The counters are copied even when the caller does not ask for the
	 * events: a run whose graph was the first work queued on the stream
	 * after the synchronisation above started it about 5 ms late on the
	 * MI355X in every 20-step run measured (with the copy queued first: 22-23
	 * us per step; without: 250 us). */
	const size_t nCounters = m_blockEvents.n;
	HIP_CHECK(hipMemcpyAsync(m_pinnedCounters, m_blockEvents.ptr, m_blockEvents.bytes(), hipMemcpyDeviceToHost,
				m_stream));
	/* a run begins with no STDP batch pending when the last thing before it
	 * was a run (which ends with its flush) */
	bool batchPending = m_stdpBlocks && m_hostStep > 0 && m_flushed != m_hostStep - 1;
	uint64_t stepsSinceApply = 0;
	const unsigned CHUNK = NEMO_STEP_RING / 2;
	/* room for the fired list at the rate of earlier runs, so that it is not
	 * reallocated (and copied) while the device waits for the host */
	const size_t firedBefore = readRaster && fired ? fired->size() : 0;
	if(readRaster && fired && counts) {
		if(fired->cap == 0) fired->spill->reserve(fired->spill->size() + size_t(1.25 * m_firedPerStep * nsteps) + 1024);
		counts->reserve(counts->size() + size_t(nsteps) * m_nets);
	}
	/* Chunks alternate between the two halves of the readout buffers (pinned
	 * bitmaps, or device and pinned lists): chunk i+1 is queued before chunk
	 * i's firing is decoded on the host, so the device does not wait for the
	 * host between chunks.  The copy of chunk i runs on its own stream beside
	 * chunk i+1; the device ring holds 128 steps, so chunk i's bitmaps (and
	 * its half's device list) are intact until chunk i+2, which waits for
	 * that copy (the copy of chunk i is the last use of its half's event
	 * before chunk i+2 records it again). */
	hipEvent_t* copied = m_readoutCopied;
	hipEvent_t* ready = m_readoutReady;
	bool inFlight[2] = {false, false};
	unsigned pendingSteps = 0; /* steps of the chunk waiting to be decoded */
	int pendingHalf = -1;
	const bool devList = readRaster && ensureListReadout();
	size_t listCopied[2] = {0, 0}; /* entries of each half's list copied with its counts */
	auto decodePending = [&]() {
		if(pendingHalf < 0) return;
		waitEvent(copied[pendingHalf]);
		if(devList) {
			const int h = pendingHalf;
			const uint32_t* cnt = m_pinnedListCounts + size_t(h) * CHUNK;
			size_t total = 0;
			for(unsigned s = 0; s < pendingSteps; ++s) {
				counts->push_back(cnt[s]);
				total += cnt[s];
			}
			if(total > listCopied[h]) {
				/* the rest, on a stream of its own: the copy stream may
				 * already wait for the next chunk */
				growPinnedList(h, total, listCopied[h]);
				HIP_CHECK(hipMemcpyAsync(m_pinnedList[h] + listCopied[h],
							m_listDev.ptr + size_t(h) * CHUNK * m_npad + listCopied[h],
							(total - listCopied[h]) * sizeof(uint32_t), hipMemcpyDeviceToHost, m_restStream));
				waitStream(m_restStream);
			}
			fired->append(m_pinnedList[h], total);
		} else {
			const uint64_t* bits = m_pinnedBits + size_t(pendingHalf) * CHUNK * m_firedWords;
			for(unsigned s = 0; s < pendingSteps; ++s) {
				decodeFiringBatch(bits + size_t(s) * m_firedWords, *fired, *counts);
			}
		}
		pendingHalf = -1;
	};
	try {
		unsigned done = 0;
		int half = 0;
		while(done < nsteps) {
			unsigned k = std::min(CHUNK, nsteps - done);
			if(stdpPeriod) k = unsigned(std::min<uint64_t>(k, stdpPeriod - stepsSinceApply));
			const uint64_t t0 = m_hostStep;
			hipGraphExec_t g = graphFor(k, batchPending);
			batchPending = true; /* each chunk leaves its steps pending */
			const bool advances = graphsAdvanceBase();
			if(!advances || m_base != t0) setBase(t0);
			/* the ring slots this chunk overwrites were copied out */
			if(inFlight[half]) HIP_CHECK(hipStreamWaitEvent(m_stream, copied[half], 0));
			HIP_CHECK(hipGraphLaunch(g, m_stream));
			if(advances) m_base = t0 + k;
			advanceSteps(k);
			if(readRaster) {
				uint32_t* listDev = devList ? m_listDev.ptr + size_t(half) * CHUNK * m_npad : nullptr;
				uint32_t* countsDev = devList ? m_listCountsDev.ptr + size_t(half) * CHUNK : nullptr;
				HIP_CHECK(hipEventRecord(ready[half], m_stream));
				HIP_CHECK(hipStreamWaitEvent(m_copyStream, ready[half], 0));
				if(devList) {
					/* compacted on the copy stream too, beside the next chunk (or
					 * the run's closing STDP batch) */
					HIP_CHECK(nemo_launch_fired_list(m_copyStream, m_firedBits.ptr, m_firedWords, t0, k,
								m_l2gDev.ptr, countsDev, listDev));
					/* the counts and as much of the list as recent rates suggest,
					 * at least 1/64 of the neurons per step (half the bytes of
					 * the bitmaps) */
					const size_t expect = std::max(size_t(1.5 * m_firedPerStep * k), size_t(k) * m_npad / 64) + 4096;
					listCopied[half] = std::min({expect, m_pinnedListCap[half], size_t(k) * m_npad});
					HIP_CHECK(hipMemcpyAsync(m_pinnedListCounts + size_t(half) * CHUNK, countsDev, k * sizeof(uint32_t),
								hipMemcpyDeviceToHost, m_copyStream));
					HIP_CHECK(hipMemcpyAsync(m_pinnedList[half], listDev, listCopied[half] * sizeof(uint32_t),
								hipMemcpyDeviceToHost, m_copyStream));
				} else {
					readBitmaps(t0, k, m_pinnedBits + size_t(half) * CHUNK * m_firedWords, m_copyStream);
				}
				HIP_CHECK(hipEventRecord(copied[half], m_copyStream));
				inFlight[half] = true;
			}
			decodePending();
			if(readRaster) {
				pendingHalf = half;
				pendingSteps = k;
				half ^= 1;
			}
			done += k;
			stepsSinceApply += k;
			if(stdpPeriod && stepsSinceApply == stdpPeriod) {
				applyStdp(reward);
				stepsSinceApply = 0;
			}
		}
		/* a run returns with the STDP accumulation of all its steps done
		 * (the last batch, or the last step's in the per-step placements) */
		enqueueFlush();
		if(events) {
			HIP_CHECK(hipMemcpyAsync(m_pinnedCounters + nCounters, m_blockEvents.ptr, m_blockEvents.bytes(),
						hipMemcpyDeviceToHost, m_stream));
		}
		decodePending();
	} catch(...) {
		(void) hipStreamSynchronize(m_stream);
		(void) hipStreamSynchronize(m_copyStream);
		throw;
	}
	waitStream(m_copyStream);
	waitStream(m_stream);
	if(readRaster && fired && nsteps) m_firedPerStep = double(fired->size() - firedBefore) / nsteps;
	if(events) {
		uint64_t before = 0, after = 0;
		for(size_t i = 0; i < nCounters; ++i) {
			before += m_pinnedCounters[i];
			after += m_pinnedCounters[nCounters + i];
		}
		*events = after - before;
	}
	m_firedOut.clear();
	checkStatus();
}

/* Eager steps with HIP events around each launch on the simulation's stream:
 * average microseconds of the propagate kernel (delivery, and STDP
 * accumulation in STDP mode 2) over the steps that have one, of the neuron
 * updates (all groups) per step, and the propagate launches per step (with
 * overlap 3 only the first step of each run of 64 has one: the others are
 * delivered by the previous update launch); and the work counters of the
 * profiled steps (events delivered, neurons fired, plastic synapses examined,
 * steps).  The steps are real simulation steps; their firing is not
 * returned. */
void Simulation::profile(unsigned nsteps, double avgUs[4], uint64_t counters[4])
{
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	const uint64_t e0 = sumCounters(m_blockEvents);
	const uint64_t x0 = sumCounters(m_blockExamined);
	std::vector<hipEvent_t> ev(4 * size_t(nsteps));
	for(hipEvent_t& e : ev) HIP_CHECK(hipEventCreate(&e));
	const uint64_t t0 = m_hostStep;
	setBase(t0);
	/* runs of up to 64 steps, as in a graph, with events between the launches */
	for(unsigned s = 0; s < nsteps; s += 64) {
		enqueueSteps(s, std::min(64u, nsteps - s), false, false, &ev[4 * size_t(s)]);
	}
	advanceSteps(nsteps);
	HIP_CHECK(hipStreamSynchronize(m_stream));
	const bool pushed = m_late.deliver_blocks && m_early.deliver_blocks && m_shape.overlap == 3;
	double sum[3] = {0, 0, 0};
	unsigned launches = 0;
	for(unsigned s = 0; s < nsteps; ++s) {
		const bool propagate = !(pushed && s % 64 != 0);
		launches += propagate ? 1 : 0;
		for(int k = propagate ? 0 : 1; k < 2; ++k) {
			float ms = 0;
			HIP_CHECK(hipEventElapsedTime(&ms, ev[4 * size_t(s) + 2 * k], ev[4 * size_t(s) + 2 * k + 1]));
			sum[k] += ms;
		}
	}
	for(hipEvent_t& e : ev) (void) hipEventDestroy(e);
	/* neurons fired in the last min(nsteps, ring) profiled steps, scaled */
	const unsigned kb = std::min<unsigned>(nsteps, NEMO_STEP_RING / 2);
	readBitmaps(m_hostStep - kb, kb);
	HIP_CHECK(hipStreamSynchronize(m_stream));
	uint64_t firedCount = 0;
	for(size_t i = 0; i < size_t(kb) * m_firedWords; ++i) firedCount += __builtin_popcountll(m_pinnedBits[i]);
	avgUs[0] = launches ? 1000.0 * sum[0] / launches : 0.0;
	avgUs[1] = nsteps ? 1000.0 * sum[1] / nsteps : 0.0;
	avgUs[2] = nsteps ? double(launches) / nsteps : 0.0;
	avgUs[3] = 0.0;
	counters[0] = sumCounters(m_blockEvents) - e0;
	counters[1] = kb ? firedCount * nsteps / kb : 0;
	counters[2] = sumCounters(m_blockExamined) - x0;
	counters[3] = nsteps;
}

void Simulation::prepare(unsigned nsteps, unsigned stdpPeriod)
{
	activate();
	uint64_t since = 0;
	unsigned done = 0;
	while(done < nsteps) {
		unsigned k = std::min(64u, nsteps - done);
		if(stdpPeriod) k = unsigned(std::min<uint64_t>(k, stdpPeriod - since));
		graphFor(k, true);
		if(done == 0) graphFor(k, false); /* a run's first chunk with nothing pending */
		done += k;
		since += k;
		if(stdpPeriod && since == stdpPeriod) since = 0;
	}
}

/* Device state a step changes (the weights change only in applyStdp). */
struct Simulation::Snapshot {
	std::vector<std::pair<void*, size_t> > regions;
	std::vector<DeviceArray<unsigned char> > copies;
	uint64_t hostStep = 0;
	uint64_t flushed = 0;
	uint64_t base = 0; /* the host's copy of the clock's base_step */
	uint64_t lastCycle = 0;
	bool updated = false;
};

void Simulation::takeSnapshot(Snapshot& s)
{
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	s.regions.clear();
	auto add = [&](void* p, size_t bytes) {
		if(p && bytes) s.regions.emplace_back(p, bytes);
	};
	for(Group& g : m_groups) {
		add(g.state.ptr, g.state.bytes());
		add(g.rng.ptr, g.rng.bytes());
	}
	add(m_clock.ptr, m_clock.bytes());
	add(m_firedBits.ptr, m_firedBits.bytes());
	add(m_acc.ptr, m_acc.bytes());
	add(m_hist.ptr, m_hist.bytes());
	add(m_stdpSnap.ptr, m_stdpSnap.bytes());
	add(m_ext.ptr, m_ext.bytes());
	add(m_rcmAcc.ptr, m_rcmAcc.bytes());
	add(m_blockEvents.ptr, m_blockEvents.bytes());
	add(m_blockExamined.ptr, m_blockExamined.bytes());
	s.copies.clear();
	s.copies.resize(s.regions.size());
	for(size_t i = 0; i < s.regions.size(); ++i) {
		s.copies[i].allocate(s.regions[i].second);
		HIP_CHECK(hipMemcpyAsync(s.copies[i].ptr, s.regions[i].first, s.regions[i].second,
					hipMemcpyDeviceToDevice, m_stream));
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	s.hostStep = m_hostStep;
	s.flushed = m_flushed;
	s.base = m_base;
	s.lastCycle = m_lastCycle;
	s.updated = m_updated;
}

void Simulation::restoreSnapshot(const Snapshot& s)
{
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	for(size_t i = 0; i < s.regions.size(); ++i) {
		HIP_CHECK(hipMemcpyAsync(s.regions[i].first, s.copies[i].ptr, s.regions[i].second,
					hipMemcpyDeviceToDevice, m_stream));
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	m_hostStep = s.hostStep;
	m_flushed = s.flushed;
	m_base = s.base;
	m_lastCycle = s.lastCycle;
	m_updated = s.updated;
}

/* Advance nsteps steps in graphs of up to 64 steps, then their pending STDP
 * accumulation (so that every STDP placement is timed with all its work);
 * seconds on the stream. */
double Simulation::timeSteps(unsigned nsteps)
{
	hipEvent_t e0, e1;
	HIP_CHECK(hipEventCreate(&e0));
	HIP_CHECK(hipEventCreate(&e1));
	HIP_CHECK(hipEventRecord(e0, m_stream));
	unsigned done = 0;
	while(done < nsteps) {
		const unsigned k = std::min(64u, nsteps - done);
		hipGraphExec_t g = graphFor(k, true);
		const bool advances = graphsAdvanceBase();
		if(!advances || m_base != m_hostStep) setBase(m_hostStep);
		HIP_CHECK(hipGraphLaunch(g, m_stream));
		if(advances) m_base = m_hostStep + k;
		advanceSteps(k);
		done += k;
	}
	enqueueFlush();
	HIP_CHECK(hipEventRecord(e1, m_stream));
	HIP_CHECK(hipStreamSynchronize(m_stream));
	float ms = 0;
	HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	return 1e-3 * ms;
}

/* Coordinate descent over delivery blocks, bin capacity, STDP fusion and
 * STDP block width; every trial runs the same steps from the same state. */
LaunchShape Simulation::autotune(unsigned stepsPerTrial)
{
	activate();
	if(m_nsynapses == 0 || stepsPerTrial == 0) return m_shape;
	stepsPerTrial = std::min(stepsPerTrial, 1024u);
	flushStdp(); /* nothing pending across shapes */
	Snapshot snap;
	takeSnapshot(snap);
	const LaunchShape start = m_shape;
	/* The snapshot holds the STDP batch clock of the shape it was taken
	 * under; a restore is therefore followed by the shape's own batch reset
	 * (applyShape does one), from the restored step with nothing pending. */
	auto trial = [&](const LaunchShape& c) -> double {
		restoreSnapshot(snap);
		try {
			applyShape(c);
		} catch(const Error&) {
			return 1e30;
		}
		timeSteps(std::min(64u, stepsPerTrial)); /* captures the graph, warms caches */
		restoreSnapshot(snap);
		resetStdpBatches();
		return timeSteps(stepsPerTrial);
	};
	LaunchShape best = start;
	double tBest = trial(best);
	auto consider = [&](const LaunchShape& c) {
		const double t = trial(c);
		if(t < tBest) {
			tBest = t;
			best = c;
		}
	};
	auto scaled = [](unsigned base, std::initializer_list<double> fs, unsigned lo, unsigned hi) {
		std::vector<unsigned> v;
		for(double f : fs) {
			const unsigned x = unsigned(std::max(double(lo), std::min(double(hi), std::round(base * f))));
			if(x != base && std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
		}
		return v;
	};
	if(m_D >= 2) {
		for(unsigned b : scaled(start.deliverBlocks, {0.25, 0.5, 2.0, 4.0}, 1, NEMO_MAX_BIN_BLOCKS - start.lateBlocks)) {
			LaunchShape c = best;
			c.deliverBlocks = b;
			consider(c);
		}
	}
	if(m_D >= 2) {
		/* splitting off the delay-1 set: a share of about 1/D of the blocks, and multiples */
		std::vector<unsigned> late;
		const unsigned base = std::max(1u, unsigned(std::lround(double(best.deliverBlocks) / m_D)));
		for(double f : {0.0, 1.0, 2.0, 4.0}) {
			const unsigned x = unsigned(std::min<double>(NEMO_MAX_BIN_BLOCKS - best.deliverBlocks, std::round(base * f)));
			if(x != best.lateBlocks && std::find(late.begin(), late.end(), x) == late.end()) late.push_back(x);
		}
		for(double f : {8.0, 16.0}) {
			const unsigned x = unsigned(std::min<double>(NEMO_MAX_BIN_BLOCKS - best.deliverBlocks, std::round(base * f)));
			if(x != best.lateBlocks && std::find(late.begin(), late.end(), x) == late.end()) late.push_back(x);
		}
		const LaunchShape before = best;
		for(unsigned b : late) {
			for(unsigned ov : {0u, 1u, 2u, 3u}) {
				if(b == 0 && ov != 0) continue;
				LaunchShape c = before;
				c.lateBlocks = b;
				c.overlap = ov;
				consider(c);
			}
		}
		if(best.lateBlocks && best.overlap >= 2) {
			/* the delays 2..D now run beside the update: retune their blocks */
			for(unsigned b : scaled(best.deliverBlocks, {0.5, 2.0}, 1, NEMO_MAX_BIN_BLOCKS - best.lateBlocks)) {
				LaunchShape c = best;
				c.deliverBlocks = b;
				consider(c);
			}
		}
	}
	for(unsigned cap : {1024u, 2048u, 4096u, 8192u, 16384u}) {
		if(cap == best.binCapacity) continue;
		LaunchShape c = best;
		c.binCapacity = cap;
		consider(c);
	}
	if(m_stdp && m_nplastic) {
		const unsigned mode0 = best.stdpMode;
		for(unsigned m : {0u, 1u, 2u, 3u, 4u}) {
			if(m == mode0) continue;
			LaunchShape c = best;
			c.stdpMode = m;
			consider(c);
		}
		for(unsigned w : {1u, 2u, 4u, 6u, 8u, 16u, 32u, 64u}) {
			if(w > (best.stdpMode == 4 ? NEMO_STDP_BATCH_MAX_WORDS : NEMO_STDP_MAX_WORDS)) continue;
			if(w == best.stdpWords) continue;
			LaunchShape c = best;
			c.stdpWords = w;
			consider(c);
		}
	}
	/* the descent keeps whatever won one trial; a shape other than the
	 * starting one is timed against it again, twice each in turn, and kept
	 * only if it is faster in total (single trials vary by a few percent) */
	auto same = [](const LaunchShape& x, const LaunchShape& y) {
		return x.deliverBlocks == y.deliverBlocks && x.lateBlocks == y.lateBlocks && x.binCapacity == y.binCapacity
			&& x.stdpMode == y.stdpMode && x.stdpWords == y.stdpWords && x.overlap == y.overlap;
	};
	if(!same(best, start)) {
		double tb = 0, ts = 0;
		for(int r = 0; r < 2; ++r) {
			tb += trial(best);
			ts += trial(start);
		}
		if(ts <= tb) best = start;
		tBest = std::min(tb, ts) / 2;
	}
	restoreSnapshot(snap);
	applyShape(best);
	if(m_logging) {
		std::printf("nemo-hip: autotuned launch shape: %u + %u delivery blocks x %u events, STDP %s, %u words, "
				"overlap %u (%.1f us/step)\n", best.deliverBlocks, best.lateBlocks, best.binCapacity,
				stdpModeName(best.stdpMode), best.stdpWords, best.overlap,
				1e6 * tBest / stepsPerTrial);
	}
	return best;
}

void Simulation::applyStdp(float reward)
{
	activate();
	if(!m_stdp) {
		throw Error(NEMO_LOGIC_ERROR, "applyStdp called, but no STDP model specified");
	}
	/* ConnectivityMatrix.cpp:279-291 */
	int32_t fx_reward = fxToFix(reward, m_fbits);
	if(fx_reward == 0 && reward != 0) {
		std::ostringstream msg;
		msg << "STDP reward rounded down to zero. The smallest valid reward is " << fxToFloat(1, m_fbits);
		throw Error(NEMO_INVALID_INPUT, msg.str());
	}
	flushStdp();
	nemo_apply_args a;
	a.rcm_of = m_rcmOf.ptr;
	a.fcm_count = m_fcm.n;
	a.rcm_acc = m_rcmAcc.ptr;
	a.fcm = m_fcm.ptr;
	a.count = m_nplastic;
	a.reward = fx_reward;
	a.minE = m_minE;
	a.maxE = m_maxE;
	a.minI = m_minI;
	a.maxI = m_maxI;
	a.fbits = m_fbits;
	HIP_CHECK(nemo_launch_apply_stdp(m_stream, &a));
	m_fcmHostValid = false;
}

/* ------------------------------------------------ externally given firing */

/* cuda/Simulation.cpp:328-346 and cpu/Simulation.cpp:115-137: the firing of
 * step s (the next step) is given; the histories and the step's bitmap are
 * written from it, the spikes arriving at step s + 1 are delivered and
 * gathered, and the timer advances by one step.  Defined only without STDP,
 * as in the CUDA backend. */
void Simulation::propagateDevice(const uint32_t* flags, float** E, float** I)
{
	activate();
	if(m_nets > 1) {
		throw Error(NEMO_API_UNSUPPORTED, "propagate drives a single network");
	}
	if(m_stdp) {
		throw Error(NEMO_LOGIC_ERROR, "Brian-specific function propagate only well-defined when STDP is not enabled");
	}
	if(!m_devToSim.ptr) {
		std::vector<uint32_t> d2s(m_npad, ~0u);
		for(size_t i = 0; i < m_ncount; ++i) d2s[m_order[i]] = uint32_t(i);
		m_devToSim.allocate(m_npad);
		HIP_CHECK(hipMemcpy(m_devToSim.ptr, d2s.data(), m_devToSim.bytes(), hipMemcpyHostToDevice));
		m_extE.allocate(std::max<size_t>(1, m_ncount));
		m_extI.allocate(std::max<size_t>(1, m_ncount));
	}
	joinStdp();
	const uint64_t s = m_hostStep;
	setBase(s);
	nemo_hip_step_buffers b = stepBuffers(0, false, false);
	b.debug_current = nullptr;
	HIP_CHECK(nemo_launch_external_fire(m_stream, &b, flags, m_devToSim.ptr, m_tiles));
	setBase(s + 1);
	launchDelivery(m_early, 0, m_stream, false);
	launchDelivery(m_late, 0, m_stream, false);
	HIP_CHECK(nemo_launch_gather_currents(m_stream, &b, m_devToSim.ptr, m_extE.ptr, m_extI.ptr, m_tiles));
	m_hostStep = s + 1;
	HIP_CHECK(hipStreamSynchronize(m_stream));
	*E = m_extE.ptr;
	*I = m_extI.ptr;
}

void Simulation::propagate(const unsigned* fired, size_t nfired, float* E, float* I, size_t n)
{
	activate();
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "current arrays must hold one entry per neuron");
	std::vector<uint32_t> flags(std::max<size_t>(1, m_ncount), 0u);
	if(nfired) {
		std::vector<uint32_t> d2s(m_npad, ~0u);
		for(size_t i = 0; i < m_ncount; ++i) d2s[m_order[i]] = uint32_t(i);
		for(size_t k = 0; k < nfired; ++k) flags[d2s[localIdx(fired[k])]] = 1u;
	}
	if(m_extFlags.n < flags.size()) m_extFlags.allocate(flags.size());
	HIP_CHECK(hipMemcpy(m_extFlags.ptr, flags.data(), flags.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
	float* dE = nullptr;
	float* dI = nullptr;
	propagateDevice(m_extFlags.ptr, &dE, &dI);
	HIP_CHECK(hipMemcpy(E, dE, n * sizeof(float), hipMemcpyDeviceToHost));
	HIP_CHECK(hipMemcpy(I, dI, n * sizeof(float), hipMemcpyDeviceToHost));
}

/* --------------------------------------------------------- neuron access */

float Simulation::getNeuronState(unsigned idx, unsigned var, unsigned net)
{
	activate();
	unsigned l = localIdx(idx, net);
	Group& g = groupOf(l);
	if(var >= g.type.nState) throw Error(NEMO_INVALID_INPUT, "Invalid state variable index " + std::to_string(var));
	float v;
	HIP_CHECK(hipStreamSynchronize(m_stream));
	HIP_CHECK(hipMemcpy(&v, statePlane(g, var) + (l - g.start), 4, hipMemcpyDeviceToHost));
	return v;
}

float Simulation::getNeuronParameter(unsigned idx, unsigned param, unsigned net)
{
	activate();
	unsigned l = localIdx(idx, net);
	Group& g = groupOf(l);
	if(param >= g.type.nParam) throw Error(NEMO_INVALID_INPUT, "Invalid parameter index " + std::to_string(param));
	float v;
	HIP_CHECK(hipStreamSynchronize(m_stream));
	HIP_CHECK(hipMemcpy(&v, g.param.ptr + size_t(param) * g.count + (l - g.start), 4, hipMemcpyDeviceToHost));
	return v;
}

float Simulation::getMembranePotential(unsigned idx, unsigned net)
{
	unsigned l = localIdx(idx, net);
	Group& g = groupOf(l);
	if(g.type.membranePotential >= g.type.nState) {
		throw Error(NEMO_INVALID_INPUT, "neuron type " + g.type.name + " has no membrane potential");
	}
	return getNeuronState(idx, g.type.membranePotential, net);
}

void Simulation::setNeuron(unsigned idx, unsigned nargs, const float* args)
{
	activate();
	unsigned l = localIdx(idx);
	Group& g = groupOf(l);
	if(nargs != g.type.nParam + g.type.nState) {
		std::ostringstream msg;
		msg << "Unexpected number of parameters/state variables when modifying neuron. Expected "
			<< (g.type.nParam + g.type.nState) << ", found " << nargs;
		throw Error(NEMO_INVALID_INPUT, msg.str());
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	for(unsigned p = 0; p < g.type.nParam; ++p) {
		HIP_CHECK(hipMemcpy(g.param.ptr + size_t(p) * g.count + (l - g.start), args + p, 4, hipMemcpyHostToDevice));
	}
	for(unsigned s = 0; s < g.type.nState; ++s) {
		HIP_CHECK(hipMemcpy(statePlane(g, s) + (l - g.start), args + g.type.nParam + s, 4,
					hipMemcpyHostToDevice));
	}
}

void Simulation::setNeuronState(unsigned idx, unsigned var, float val)
{
	activate();
	unsigned l = localIdx(idx);
	Group& g = groupOf(l);
	if(var >= g.type.nState) throw Error(NEMO_INVALID_INPUT, "Invalid state variable index " + std::to_string(var));
	HIP_CHECK(hipStreamSynchronize(m_stream));
	HIP_CHECK(hipMemcpy(statePlane(g, var) + (l - g.start), &val, 4, hipMemcpyHostToDevice));
}

void Simulation::setNeuronParameter(unsigned idx, unsigned param, float val)
{
	activate();
	unsigned l = localIdx(idx);
	Group& g = groupOf(l);
	if(param >= g.type.nParam) throw Error(NEMO_INVALID_INPUT, "Invalid parameter index " + std::to_string(param));
	HIP_CHECK(hipStreamSynchronize(m_stream));
	HIP_CHECK(hipMemcpy(g.param.ptr + size_t(param) * g.count + (l - g.start), &val, 4, hipMemcpyHostToDevice));
}

/* Vector forms of the neuron getters and setters (the Python binding's list
 * arguments, src/api/python/nemo_py.cpp): a few neurons are read or written
 * element by element; more, plane by plane (each plane of a model that the
 * call touches is copied to the host once, and written back once). */
static const size_t ELEMENTWISE_MAX = 16;

void Simulation::getNeuronValues(bool param, unsigned var, size_t n, const unsigned* idx, float* out)
{
	activate();
	std::vector<unsigned> loc(n);
	for(size_t i = 0; i < n; ++i) {
		loc[i] = localIdx(idx[i]);
		const Group& g = groupOf(loc[i]);
		if(var >= (param ? g.type.nParam : g.type.nState)) {
			throw Error(NEMO_INVALID_INPUT, std::string(param ? "Invalid parameter index " : "Invalid state variable index ")
					+ std::to_string(var));
		}
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	std::vector<std::vector<float> > planes(m_groups.size());
	for(size_t i = 0; i < n; ++i) {
		const int gi = m_typeOfLocal[loc[i]];
		Group& g = m_groups[gi];
		const float* plane = (param ? g.param.ptr + size_t(var) * g.count : statePlane(g, var));
		if(n <= ELEMENTWISE_MAX) {
			HIP_CHECK(hipMemcpy(out + i, plane + (loc[i] - g.start), 4, hipMemcpyDeviceToHost));
			continue;
		}
		if(planes[gi].empty()) {
			planes[gi].resize(g.count);
			HIP_CHECK(hipMemcpy(planes[gi].data(), plane, g.count * sizeof(float), hipMemcpyDeviceToHost));
		}
		out[i] = planes[gi][loc[i] - g.start];
	}
}

void Simulation::setNeuronValues(bool param, unsigned var, size_t n, const unsigned* idx, const float* vals)
{
	activate();
	std::vector<unsigned> loc(n);
	for(size_t i = 0; i < n; ++i) {
		loc[i] = localIdx(idx[i]);
		const Group& g = groupOf(loc[i]);
		if(var >= (param ? g.type.nParam : g.type.nState)) {
			throw Error(NEMO_INVALID_INPUT, std::string(param ? "Invalid parameter index " : "Invalid state variable index ")
					+ std::to_string(var));
		}
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	if(n <= ELEMENTWISE_MAX) {
		for(size_t i = 0; i < n; ++i) {
			Group& g = groupOf(loc[i]);
			float* plane = (param ? g.param.ptr + size_t(var) * g.count : statePlane(g, var));
			HIP_CHECK(hipMemcpy(plane + (loc[i] - g.start), vals + i, 4, hipMemcpyHostToDevice));
		}
		return;
	}
	std::vector<std::vector<float> > planes(m_groups.size());
	for(size_t i = 0; i < n; ++i) {
		const int gi = m_typeOfLocal[loc[i]];
		Group& g = m_groups[gi];
		if(planes[gi].empty()) {
			planes[gi].resize(g.count);
			HIP_CHECK(hipMemcpy(planes[gi].data(), (param ? g.param.ptr + size_t(var) * g.count : statePlane(g, var)),
						g.count * sizeof(float), hipMemcpyDeviceToHost));
		}
		planes[gi][loc[i] - g.start] = vals[i]; /* for repeated neurons the last value wins */
	}
	for(size_t gi = 0; gi < m_groups.size(); ++gi) {
		if(planes[gi].empty()) continue;
		Group& g = m_groups[gi];
		HIP_CHECK(hipMemcpy((param ? g.param.ptr + size_t(var) * g.count : statePlane(g, var)), planes[gi].data(),
					g.count * sizeof(float), hipMemcpyHostToDevice));
	}
}

void Simulation::setNeurons(size_t n, const unsigned* idx, unsigned nargs, const float* args)
{
	activate();
	std::vector<unsigned> loc(n);
	for(size_t i = 0; i < n; ++i) {
		loc[i] = localIdx(idx[i]);
		const Group& g = groupOf(loc[i]);
		if(nargs != g.type.nParam + g.type.nState) {
			std::ostringstream msg;
			msg << "Unexpected number of parameters/state variables when modifying neuron. Expected "
				<< (g.type.nParam + g.type.nState) << ", found " << nargs;
			throw Error(NEMO_INVALID_INPUT, msg.str());
		}
	}
	if(n <= ELEMENTWISE_MAX) {
		for(size_t i = 0; i < n; ++i) setNeuron(idx[i], nargs, args + i * nargs);
		return;
	}
	HIP_CHECK(hipStreamSynchronize(m_stream));
	/* per touched group: all its parameter and state planes, [plane][i] */
	std::vector<std::vector<float> > param(m_groups.size()), state(m_groups.size());
	for(size_t i = 0; i < n; ++i) {
		const int gi = m_typeOfLocal[loc[i]];
		Group& g = m_groups[gi];
		const size_t np = size_t(g.type.nParam) * g.count, ns = size_t(g.type.nState) * g.count;
		if(param[gi].empty() && state[gi].empty()) {
			param[gi].resize(np);
			state[gi].resize(ns);
			if(np) HIP_CHECK(hipMemcpy(param[gi].data(), g.param.ptr, np * sizeof(float), hipMemcpyDeviceToHost));
			if(ns) HIP_CHECK(hipMemcpy(state[gi].data(), statePlane(g, 0), ns * sizeof(float), hipMemcpyDeviceToHost));
			if(param[gi].empty() && state[gi].empty()) continue; /* a model without parameters or state */
		}
		const size_t k = loc[i] - g.start;
		for(unsigned p = 0; p < g.type.nParam; ++p) param[gi][size_t(p) * g.count + k] = args[i * nargs + p];
		for(unsigned v = 0; v < g.type.nState; ++v) state[gi][size_t(v) * g.count + k] = args[i * nargs + g.type.nParam + v];
	}
	for(size_t gi = 0; gi < m_groups.size(); ++gi) {
		Group& g = m_groups[gi];
		if(!param[gi].empty()) {
			HIP_CHECK(hipMemcpy(g.param.ptr, param[gi].data(), param[gi].size() * sizeof(float), hipMemcpyHostToDevice));
		}
		if(!state[gi].empty()) {
			HIP_CHECK(hipMemcpy(statePlane(g, 0), state[gi].data(), state[gi].size() * sizeof(float),
						hipMemcpyHostToDevice));
		}
	}
}

/* -------------------------------------------------------- synapse access */

void Simulation::syncFcmHost()
{
	if(m_fcmHostValid) return;
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	m_fcmHost.resize(m_fcm.n);
	HIP_CHECK(hipMemcpy(m_fcmHost.data(), m_fcm.ptr, m_fcm.bytes(), hipMemcpyDeviceToHost));
	m_fcmHostValid = true;
}

size_t Simulation::synapseIndex(synapse_id id) const
{
	if(m_writeOnly) {
		throw Error(NEMO_INVALID_INPUT, "Cannot read synapse state if simulation configured with write-only synapses");
	}
	unsigned src = unsigned(id >> 32);
	size_t k = size_t(id & 0xffffffffu);
	auto it = m_synRange.find(netKey(0, src));
	if(it == m_synRange.end()) {
		throw Error(NEMO_INVALID_INPUT, "Non-existing neuron id (" + std::to_string(src) + ") in synapse query");
	}
	if(k >= it->second.second - it->second.first) {
		throw Error(NEMO_INVALID_INPUT, "synapse id out of range");
	}
	return it->second.first + k;
}

const std::vector<synapse_id>& Simulation::getSynapsesFrom(unsigned source)
{
	if(m_writeOnly) {
		throw Error(NEMO_INVALID_INPUT, "Cannot read synapse state if simulation configured with write-only synapses");
	}
	m_queried.clear();
	auto it = m_synRange.find(netKey(0, source));
	if(it == m_synRange.end()) {
		localIdx(source); /* throws for a non-existing neuron */
	} else {
		const size_t n = it->second.second - it->second.first;
		m_queried.resize(n);
		for(size_t k = 0; k < n; ++k) m_queried[k] = (synapse_id(source) << 32) | synapse_id(k);
	}
	return m_queried;
}

unsigned Simulation::getSynapseTarget(synapse_id id) { return m_synTarget[synapseIndex(id)]; }
unsigned Simulation::getSynapseDelay(synapse_id id) { return m_synDelay[synapseIndex(id)]; }
unsigned char Simulation::getSynapsePlastic(synapse_id id) { return m_synPlastic[synapseIndex(id)]; }

float Simulation::getSynapseWeight(synapse_id id)
{
	size_t s = synapseIndex(id);
	syncFcmHost();
	return fxToFloat(m_fcmHost[m_synAddr[s]].y, m_fbits);
}

/* ----------------------------------------------------------------- timer */

unsigned long Simulation::elapsedWallclock() const
{
	using namespace std::chrono;
	(void) hipStreamSynchronize(m_stream);
	return (unsigned long) duration_cast<milliseconds>(steady_clock::now() - m_wallStart).count();
}

unsigned long Simulation::elapsedSimulation()
{
	return (unsigned long) (m_hostStep - m_cycleOrigin);
}

void Simulation::resetTimer()
{
	activate();
	m_cycleOrigin = m_hostStep;
	copyScalar(&m_clock.ptr->cycle_origin, m_cycleOrigin);
	HIP_CHECK(hipStreamSynchronize(m_stream));
	m_wallStart = std::chrono::steady_clock::now();
}

/* ------------------------------------------------------------- inspection */

void Simulation::setDebug(bool on)
{
	m_debug = on;
}

void Simulation::getCurrents(float* E, float* I, size_t n)
{
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	std::vector<float> h(2 * m_accPitch);
	HIP_CHECK(hipMemcpy(h.data(), m_debugCurrent.ptr, h.size() * 4, hipMemcpyDeviceToHost));
	for(size_t i = 0; i < n; ++i) {
		E[i] = h[m_order[i]];
		I[i] = h[m_accPitch + m_order[i]];
	}
}

void Simulation::localToGlobal(unsigned* g, size_t n) const
{
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	for(size_t i = 0; i < n; ++i) g[i] = m_l2g[m_order[i]];
}

void Simulation::getRecentFiring(uint64_t* h, size_t n)
{
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	std::vector<uint64_t> d(m_npad);
	/* the history words after the last step */
	const uint64_t* last = m_hist.ptr + NEMO_HIST_INDEX(0, m_hostStep == 0 ? 1 : m_hostStep - 1, m_accPitch);
	HIP_CHECK(hipMemcpy(d.data(), last, m_npad * 8, hipMemcpyDeviceToHost));
	for(size_t i = 0; i < n; ++i) h[i] = d[m_order[i]];
}

void Simulation::getRngState(uint32_t* s, size_t n)
{
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	for(Group& g : m_groups) {
		std::vector<uint32_t> h(g.rng.n);
		HIP_CHECK(hipMemcpy(h.data(), g.rng.ptr, g.rng.bytes(), hipMemcpyDeviceToHost));
		for(size_t i = 0; i < n; ++i) {
			const unsigned l = m_order[i];
			if(l >= g.start && l < g.start + g.count) {
				std::copy(h.begin() + 4 * size_t(l - g.start), h.begin() + 4 * size_t(l - g.start + 1), s + 4 * i);
			}
		}
	}
}

void Simulation::getNeuronStates(unsigned var, float* out, size_t n)
{
	if(n != m_ncount) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	std::fill(out, out + n, std::nanf(""));
	for(Group& g : m_groups) {
		if(var >= g.type.nState || g.count == 0) continue;
		std::vector<float> h(g.count);
		HIP_CHECK(hipMemcpy(h.data(), statePlane(g, var), g.count * sizeof(float),
					hipMemcpyDeviceToHost));
		for(size_t i = 0; i < n; ++i) {
			const unsigned l = m_order[i];
			if(l >= g.start && l < g.start + g.count) out[i] = h[l - g.start];
		}
	}
}

void Simulation::getFixedWeights(int32_t* w, size_t n)
{
	if(m_writeOnly) throw Error(NEMO_INVALID_INPUT, "write-only synapses");
	if(n != m_nsynapses) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	syncFcmHost();
	for(size_t k = 0; k < n; ++k) w[k] = m_fcmHost[m_synAddr[k]].y;
}

void Simulation::getStdpAccumulators(int32_t* acc, size_t n)
{
	if(m_writeOnly) throw Error(NEMO_INVALID_INPUT, "write-only synapses");
	if(n != m_nsynapses) throw Error(NEMO_INVALID_INPUT, "size mismatch");
	std::fill(acc, acc + n, 0);
	if(!m_stdp || m_nplastic == 0) return;
	activate();
	flushStdp();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	std::vector<int32_t> racc(m_nplastic);
	std::vector<uint32_t> fwd(m_nplastic);
	HIP_CHECK(hipMemcpy(racc.data(), m_rcmAcc.ptr, m_nplastic * 4, hipMemcpyDeviceToHost));
	HIP_CHECK(hipMemcpy(fwd.data(), m_rcmFwd.ptr, m_nplastic * 4, hipMemcpyDeviceToHost));
	std::vector<int32_t> byAddr(m_fcm.n, 0);
	for(size_t r = 0; r < m_nplastic; ++r) byAddr[fwd[r]] = racc[r];
	for(size_t k = 0; k < n; ++k) acc[k] = byAddr[m_synAddr[k]];
}

uint64_t Simulation::eventsDelivered()
{
	activate();
	HIP_CHECK(hipStreamSynchronize(m_stream));
	return sumCounters(m_blockEvents);
}

} // namespace nemo_hip
