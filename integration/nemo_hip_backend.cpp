/*
 * nemo_hip_backend.cpp -- the NeMo side of the HIP backend.
 *
 * Implements nemo::SimulationBackend (src/nemo/SimulationBackend.hpp:39-163,
 * with the nemo::Simulation and nemo::ReadableNetwork queries it inherits)
 * over the C ABI of libnemo_hip.so (include/nemo_hip.h), and exports the
 * backend factory that NeMo's dispatch resolves by name, with the signatures
 * of the CUDA backend's:
 *
 *   hip_simulation          src/nemo/cuda/create_simulation.hpp:14-22
 *   hip_set_configuration   src/nemo/cuda/devices.hpp:23-30
 *   hip_test_device         src/nemo/cuda/devices.hpp:33-37
 *   hip_device_count        src/nemo/cuda/devices.hpp:40-44
 *   hip_device_description  src/nemo/cuda/devices.hpp:47-51
 *
 * This file is compiled inside NeMo's source tree, against NeMo's headers
 * and libnemo_base, into the plugin library libnemo_hip_backend; the
 * dispatch case NeMo needs for it is shown beside the build rule in the
 * integration notes.  It is not compiled in this repository: NeMo's headers
 * need Boost and CMake-generated files that are absent here.
 *
 * NeMo's libnemo exports the same nemo_* C names as libnemo_hip.so, so this
 * adapter does not link against libnemo_hip.so: it opens it privately
 * (RTLD_LOCAL) and resolves the entry points it uses into a table; the
 * library binds its own calls to itself (it is linked -Bsymbolic).
 * NEMO_HIP_LIBRARY names the library; by default it is found beside this
 * plugin.
 */
#include <dlfcn.h>
#include <stdint.h>

#include <cstdlib>
#include <string>
#include <vector>

#include <nemo/config.h>
#include <nemo/ConfigurationImpl.hpp>
#include <nemo/FiringBuffer.hpp>
#include <nemo/NeuronType.hpp>
#include <nemo/SimulationBackend.hpp>
#include <nemo/exception.hpp>
#include <nemo/network/Generator.hpp>

/* the factory's symbols stay visible whatever visibility the plugin is
 * built with */
#define HIP_BACKEND_PUBLIC __attribute__((visibility("default")))

namespace {

/* The entry points of include/nemo_hip.h the adapter calls; handles are
 * opaque. */
struct HipApi
{
	void* handle;
	const char* (*strerror)(void);
	int (*device_count)(unsigned*);
	int (*device_description)(unsigned, const char**);
	int (*test_device)(unsigned);
	void* (*new_configuration)(void);
	void (*delete_configuration)(void*);
	int (*log_stdout)(void*);
	int (*set_stdp_function)(void*, float*, size_t, float*, size_t, float, float, float, float);
	int (*set_hip_backend)(void*, int);
	int (*set_write_only_synapses)(void*);
	void* (*new_network)(void);
	void (*delete_network)(void*);
	int (*add_neuron_type)(void*, const char*, unsigned*);
	int (*add_neurons)(void*, unsigned, size_t, const unsigned*, unsigned, const float*);
	int (*add_synapses)(void*, size_t, const unsigned*, const unsigned*, const unsigned*, const float*,
			const unsigned char*, uint64_t*);
	void* (*new_simulation)(void*, void*);
	void (*delete_simulation)(void*);
	int (*step)(void*, unsigned*, size_t, unsigned*, float*, size_t, unsigned**, size_t*);
	int (*apply_stdp)(void*, float);
	int (*get_membrane_potential)(void*, unsigned, float*);
	int (*get_neuron_state)(void*, unsigned, unsigned, float*);
	int (*get_neuron_parameter)(void*, unsigned, unsigned, float*);
	int (*set_neuron)(void*, unsigned, unsigned, float*);
	int (*set_neuron_state)(void*, unsigned, unsigned, float);
	int (*set_neuron_parameter)(void*, unsigned, unsigned, float);
	int (*get_synapses_from)(void*, unsigned, uint64_t**, size_t*);
	int (*get_synapse_target)(void*, uint64_t, unsigned*);
	int (*get_synapse_delay)(void*, uint64_t, unsigned*);
	int (*get_synapse_weight)(void*, uint64_t, float*);
	int (*get_synapse_plastic)(void*, uint64_t, unsigned char*);
	int (*elapsed_wallclock)(void*, unsigned long*);
	int (*elapsed_simulation)(void*, unsigned long*);
	int (*reset_timer)(void*);
	int (*counts)(void*, size_t*, size_t*, size_t*);
	int (*local_to_global)(void*, unsigned*, size_t);
	int (*propagate_device)(void*, const uint32_t*, float**, float**);
};

template<typename F>
void resolve(void* h, const char* name, F& fn)
{
	fn = reinterpret_cast<F>(dlsym(h, name));
	if(!fn) {
		throw nemo::exception(NEMO_DL_ERROR, std::string("libnemo_hip.so does not export ") + name);
	}
}

std::string pluginDirectory()
{
	Dl_info info;
	if(dladdr(reinterpret_cast<void*>(&pluginDirectory), &info) && info.dli_fname) {
		std::string p(info.dli_fname);
		size_t slash = p.rfind('/');
		if(slash != std::string::npos) return p.substr(0, slash);
	}
	return ".";
}

const HipApi& api()
{
	static HipApi a = [] {
		HipApi t;
		const char* env = std::getenv("NEMO_HIP_LIBRARY");
		const std::string path = env ? std::string(env) : pluginDirectory() + "/libnemo_hip.so";
		t.handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
		if(!t.handle) {
			throw nemo::exception(NEMO_DL_ERROR, std::string("failed to load ") + path + ": " + dlerror());
		}
		void* h = t.handle;
		resolve(h, "nemo_strerror", t.strerror);
		resolve(h, "nemo_hip_device_count", t.device_count);
		resolve(h, "nemo_hip_device_description", t.device_description);
		resolve(h, "nemo_hip_test_device", t.test_device);
		resolve(h, "nemo_new_configuration", t.new_configuration);
		resolve(h, "nemo_delete_configuration", t.delete_configuration);
		resolve(h, "nemo_log_stdout", t.log_stdout);
		resolve(h, "nemo_set_stdp_function", t.set_stdp_function);
		resolve(h, "nemo_set_hip_backend", t.set_hip_backend);
		resolve(h, "nemo_set_write_only_synapses", t.set_write_only_synapses);
		resolve(h, "nemo_new_network", t.new_network);
		resolve(h, "nemo_delete_network", t.delete_network);
		resolve(h, "nemo_add_neuron_type", t.add_neuron_type);
		resolve(h, "nemo_add_neurons", t.add_neurons);
		resolve(h, "nemo_add_synapses", t.add_synapses);
		resolve(h, "nemo_new_simulation", t.new_simulation);
		resolve(h, "nemo_delete_simulation", t.delete_simulation);
		resolve(h, "nemo_step", t.step);
		resolve(h, "nemo_apply_stdp", t.apply_stdp);
		resolve(h, "nemo_get_membrane_potential", t.get_membrane_potential);
		resolve(h, "nemo_get_neuron_state_s", t.get_neuron_state);
		resolve(h, "nemo_get_neuron_parameter_s", t.get_neuron_parameter);
		resolve(h, "nemo_set_neuron_s", t.set_neuron);
		resolve(h, "nemo_set_neuron_state_s", t.set_neuron_state);
		resolve(h, "nemo_set_neuron_parameter_s", t.set_neuron_parameter);
		resolve(h, "nemo_get_synapses_from_s", t.get_synapses_from);
		resolve(h, "nemo_get_synapse_target_s", t.get_synapse_target);
		resolve(h, "nemo_get_synapse_delay_s", t.get_synapse_delay);
		resolve(h, "nemo_get_synapse_weight_s", t.get_synapse_weight);
		resolve(h, "nemo_get_synapse_plastic_s", t.get_synapse_plastic);
		resolve(h, "nemo_elapsed_wallclock", t.elapsed_wallclock);
		resolve(h, "nemo_elapsed_simulation", t.elapsed_simulation);
		resolve(h, "nemo_reset_timer", t.reset_timer);
		resolve(h, "nemo_hip_counts", t.counts);
		resolve(h, "nemo_hip_local_to_global", t.local_to_global);
		resolve(h, "nemo_hip_propagate_device", t.propagate_device);
		return t;
	}();
	return a;
}

/* A C-ABI status as a nemo::exception (nemo_c.cpp:20-60 in reverse) */
void call(int status)
{
	if(status != NEMO_OK) {
		throw nemo::exception(status, api().strerror());
	}
}

/* Owning wrappers of the C handles */
struct Handle
{
	void* p;
	void (*release)(void*);
	Handle(void* p_, void (*r)(void*)) : p(p_), release(r) {}
	~Handle() { if(p) release(p); }
	Handle(const Handle&) = delete;
	Handle& operator=(const Handle&) = delete;
};


/* The configuration the HIP backend needs from NeMo's: STDP function,
 * write-only synapses, logging, device (ConfigurationImpl.hpp:42-107) */
void* translateConfiguration(const nemo::ConfigurationImpl& conf)
{
	const HipApi& a = api();
	void* c = a.new_configuration();
	try {
		if(conf.fractionalBitsSet() && conf.fractionalBits() != 20) {
			throw nemo::exception(NEMO_API_UNSUPPORTED, "the HIP backend computes currents in Q11.20 only");
		}
		const boost::optional<nemo::StdpFunction>& fn = conf.stdpFunction();
		if(fn) {
			std::vector<float> pre(fn->prefire()), post(fn->postfire());
			call(a.set_stdp_function(c, pre.empty() ? NULL : &pre[0], pre.size(),
						post.empty() ? NULL : &post[0], post.size(),
						fn->minExcitatoryWeight(), fn->maxExcitatoryWeight(),
						fn->minInhibitoryWeight(), fn->maxInhibitoryWeight()));
		}
		if(conf.writeOnlySynapses()) call(a.set_write_only_synapses(c));
		if(conf.loggingEnabled()) call(a.log_stdout(c));
		call(a.set_hip_backend(c, int(conf.cudaDevice())));
	} catch(...) {
		a.delete_configuration(c);
		throw;
	}
	return c;
}


/* The network of a Generator (neurons type by type, then synapses in source
 * order with their ids, network/Generator.hpp:17-61) in the library's
 * Network.  Synapse ids are assigned by the library as (source << 32) |
 * index in insertion order per source, which is the id order the
 * programmatic generator iterates in (network/programmatic/
 * synapse_iterator.hpp); any other order is refused. */
void* translateNetwork(const nemo::network::Generator& net)
{
	const HipApi& a = api();
	void* n = a.new_network();
	try {
		for(unsigned t = 0; t < net.neuronTypeCount(); ++t) {
			const nemo::NeuronType& type = net.neuronType(t);
			unsigned id = 0;
			call(a.add_neuron_type(n, type.name().c_str(), &id));
			const unsigned nargs = unsigned(type.parameterCount() + type.stateVarCount());
			std::vector<unsigned> idx;
			std::vector<float> args;
			for(nemo::network::neuron_iterator i = net.neuron_begin(t), end = net.neuron_end(t); i != end; ++i) {
				idx.push_back(i->first);
				const nemo::Neuron& neuron = i->second;
				for(size_t p = 0; p < type.parameterCount(); ++p) args.push_back(neuron.getParameter(p));
				for(size_t s = 0; s < type.stateVarCount(); ++s) args.push_back(neuron.getState(s));
			}
			if(!idx.empty()) {
				call(a.add_neurons(n, id, idx.size(), &idx[0], nargs, args.empty() ? NULL : &args[0]));
			}
		}
		const size_t CHUNK = size_t(1) << 20;
		std::vector<unsigned> src, tgt, delay;
		std::vector<float> weight;
		std::vector<unsigned char> plastic;
		std::vector<uint64_t> want;
		std::vector<uint64_t> got;
		auto flush = [&]() {
			if(src.empty()) return;
			got.resize(src.size());
			call(a.add_synapses(n, src.size(), &src[0], &tgt[0], &delay[0], &weight[0], &plastic[0], &got[0]));
			if(got != want) {
				throw nemo::exception(NEMO_LOGIC_ERROR,
						"the HIP backend needs the synapses of each source in the order of their ids");
			}
			src.clear(); tgt.clear(); delay.clear(); weight.clear(); plastic.clear(); want.clear();
		};
		for(nemo::network::synapse_iterator i = net.synapse_begin(), end = net.synapse_end(); i != end; ++i) {
			src.push_back(i->source);
			tgt.push_back(i->target());
			delay.push_back(i->delay);
			weight.push_back(i->weight());
			plastic.push_back(i->plastic());
			want.push_back((uint64_t(i->source) << 32) | uint64_t(i->id()));
			if(src.size() == CHUNK) flush();
		}
		flush();
	} catch(...) {
		a.delete_network(n);
		throw;
	}
	return n;
}


/* The simulation object NeMo drives through SimulationBackend: the step's
 * stimuli are collected in prefire/set*Stimulus and the library's step runs
 * in fire (delivery, neuron update, STDP); postfire has nothing left to do
 * and readFiring returns the step's fired neurons, which stay valid until
 * the next step (Simulation.hpp:56-63). */
class HipSimulation : public nemo::SimulationBackend
{
	public :

		HipSimulation(const nemo::network::Generator& net, const nemo::ConfigurationImpl& conf) :
			m_conf(translateConfiguration(conf), api().delete_configuration),
			m_net(translateNetwork(net), api().delete_network),
			m_sim(NULL, api().delete_simulation),
			m_cycle(0)
		{
			m_sim.p = api().new_simulation(m_net.p, m_conf.p);
			if(!m_sim.p) {
				throw nemo::exception(NEMO_CUDA_ERROR, api().strerror());
			}
			size_t n = 0, s = 0, p = 0;
			call(api().counts(m_sim.p, &n, &s, &p));
			m_l2g.resize(n);
			call(api().local_to_global(m_sim.p, n ? &m_l2g[0] : NULL, n));
		}

		unsigned getFractionalBits() const { return 20; }

		void setFiringStimulus(const std::vector<unsigned>& nidx) { m_fstim = nidx; }

		void initCurrentStimulus(size_t count)
		{
			m_istimIdx.clear();
			m_istimVal.clear();
			m_istimIdx.reserve(count);
			m_istimVal.reserve(count);
		}

		void addCurrentStimulus(unsigned neuron, float current)
		{
			m_istimIdx.push_back(neuron);
			m_istimVal.push_back(current);
		}

		void finalizeCurrentStimulus(size_t) { }

		/* current for every neuron, in the backend's own (simulation) order */
		void setCurrentStimulus(const std::vector<float>& current)
		{
			if(current.size() != m_l2g.size()) {
				throw nemo::exception(NEMO_INVALID_INPUT, "current stimulus vector does not match the network size");
			}
			initCurrentStimulus(current.size());
			for(size_t i = 0; i < current.size(); ++i) {
				if(current[i] != 0.0f) addCurrentStimulus(m_l2g[i], current[i]);
			}
		}

		void prefire() { }

		void fire()
		{
			unsigned* fired = NULL;
			size_t nfired = 0;
			call(api().step(m_sim.p, m_fstim.empty() ? NULL : &m_fstim[0], m_fstim.size(),
						m_istimIdx.empty() ? NULL : &m_istimIdx[0],
						m_istimVal.empty() ? NULL : &m_istimVal[0], m_istimIdx.size(), &fired, &nfired));
			m_fired.assign(fired, fired + nfired);
			m_fstim.clear();
			m_istimIdx.clear();
			m_istimVal.clear();
			m_cycle += 1;
		}

		void postfire() { }

		nemo::FiredList readFiring() { return nemo::FiredList(m_cycle - 1, m_fired); }

		void applyStdp(float reward) { call(api().apply_stdp(m_sim.p, reward)); }

		void setNeuron(unsigned idx, unsigned nargs, const float args[])
		{
			std::vector<float> a(args, args + nargs);
			call(api().set_neuron(m_sim.p, idx, nargs, a.empty() ? NULL : &a[0]));
		}

		void setNeuronState(unsigned neuron, unsigned var, float value)
		{
			call(api().set_neuron_state(m_sim.p, neuron, var, value));
		}

		void setNeuronParameter(unsigned neuron, unsigned parameter, float value)
		{
			call(api().set_neuron_parameter(m_sim.p, neuron, parameter, value));
		}

		float getNeuronState(unsigned neuron, unsigned var) const
		{
			float v = 0.0f;
			call(api().get_neuron_state(m_sim.p, neuron, var, &v));
			return v;
		}

		float getNeuronParameter(unsigned neuron, unsigned parameter) const
		{
			float v = 0.0f;
			call(api().get_neuron_parameter(m_sim.p, neuron, parameter, &v));
			return v;
		}

		float getMembranePotential(unsigned neuron) const
		{
			float v = 0.0f;
			call(api().get_membrane_potential(m_sim.p, neuron, &v));
			return v;
		}

		const std::vector<synapse_id>& getSynapsesFrom(unsigned neuron)
		{
			uint64_t* ids = NULL;
			size_t len = 0;
			call(api().get_synapses_from(m_sim.p, neuron, &ids, &len));
			m_queried.assign(ids, ids + len);
			return m_queried;
		}

		unsigned getSynapseTarget(const synapse_id& s) const
		{
			unsigned v = 0;
			call(api().get_synapse_target(m_sim.p, s, &v));
			return v;
		}

		unsigned getSynapseDelay(const synapse_id& s) const
		{
			unsigned v = 0;
			call(api().get_synapse_delay(m_sim.p, s, &v));
			return v;
		}

		float getSynapseWeight(const synapse_id& s) const
		{
			float v = 0.0f;
			call(api().get_synapse_weight(m_sim.p, s, &v));
			return v;
		}

		unsigned char getSynapsePlastic(const synapse_id& s) const
		{
			unsigned char v = 0;
			call(api().get_synapse_plastic(m_sim.p, s, &v));
			return v;
		}

		unsigned long elapsedWallclock() const
		{
			unsigned long v = 0;
			call(api().elapsed_wallclock(m_sim.p, &v));
			return v;
		}

		unsigned long elapsedSimulation() const
		{
			unsigned long v = 0;
			call(api().elapsed_simulation(m_sim.p, &v));
			return v;
		}

		void resetTimer() { call(api().reset_timer(m_sim.p)); }

#ifdef NEMO_BRIAN_ENABLED
		/* cuda/Simulation.cpp:328-346: fired is a device array of one flag per
		 * neuron (simulation order); the returned device arrays hold the E
		 * and I currents of the next step, valid until the next call */
		std::pair<float*, float*> propagate_raw(uint32_t* fired, int /* nfired */)
		{
			float* E = NULL;
			float* I = NULL;
			call(api().propagate_device(m_sim.p, fired, &E, &I));
			m_cycle += 1;
			return std::make_pair(E, I);
		}
#endif

	private :

		Handle m_conf, m_net, m_sim;
		std::vector<unsigned> m_l2g;   /* simulation order -> neuron index */
		std::vector<unsigned> m_fstim;
		std::vector<unsigned> m_istimIdx;
		std::vector<float> m_istimVal;
		std::vector<unsigned> m_fired;
		std::vector<synapse_id> m_queried;
		nemo::cycle_t m_cycle;
};

} // namespace


extern "C" {

/* create_simulation.hpp:14-22 */
HIP_BACKEND_PUBLIC
nemo::SimulationBackend*
hip_simulation(const nemo::network::Generator* net, const nemo::ConfigurationImpl* conf)
{
	return new HipSimulation(*net, *conf);
}

/* devices.hpp:23-30 and devices.cpp:226-244: choose the device (-1: the
 * default) and record the choice in the configuration */
HIP_BACKEND_PUBLIC
void
hip_set_configuration(nemo::ConfigurationImpl* conf, int dev)
{
	unsigned count = 0;
	call(api().device_count(&count));
	if(count == 0) {
		throw nemo::exception(NEMO_CUDA_ERROR, "No HIP devices available");
	}
	if(dev >= int(count)) {
		throw nemo::exception(NEMO_CUDA_ERROR, "Invalid device: " + std::to_string(dev));
	}
	const unsigned device = dev < 0 ? 0u : unsigned(dev);
	conf->setBackend(NEMO_BACKEND_HIP);
	conf->setCudaDevice(device);
	const char* descr = NULL;
	call(api().device_description(device, &descr));
	conf->setBackendDescription(descr);
}

/* devices.hpp:33-37: throws if the device cannot run the backend */
HIP_BACKEND_PUBLIC
void
hip_test_device(unsigned device)
{
	call(api().test_device(device));
}

/* devices.hpp:40-44 */
HIP_BACKEND_PUBLIC
unsigned
hip_device_count(void)
{
	unsigned n = 0;
	return api().device_count(&n) == NEMO_OK ? n : 0u;
}

/* devices.hpp:47-51 */
HIP_BACKEND_PUBLIC
const char*
hip_device_description(unsigned device)
{
	const char* d = NULL;
	call(api().device_description(device, &d));
	return d;
}

}
