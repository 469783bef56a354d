/*
 * C ABI of libnemo_hip.so (include/nemo_hip.h).  Mirrors NeMo's C API
 * (src/nemo/nemo_c.cpp): exceptions are caught at the boundary and turned into
 * a status code plus a last-error string (nemo_c.cpp:20-60).
 */
#include <cstring>
#include <string>
#include <vector>

#include "backend.hpp"

using nemo_hip::Error;

struct nemo_network_class { nemo_hip::Network impl; };
struct nemo_configuration_class { nemo_hip::Configuration impl; };
struct nemo_simulation_class {
	explicit nemo_simulation_class(const nemo_hip::Network& n, const nemo_hip::Configuration& c) : impl(n, c) {}
	nemo_simulation_class(const std::vector<const nemo_hip::Network*>& n, const nemo_hip::Configuration& c)
		: impl(n, c) {}
	nemo_hip::Simulation impl;
	std::vector<unsigned> runCounts, runFired;
};

namespace {

thread_local std::string g_lastError = "No error";
thread_local std::string g_descr;

template<typename F>
nemo_status_t guarded(F f)
{
	try {
		f();
		return NEMO_OK;
	} catch(Error& e) {
		g_lastError = e.what();
		return e.code;
	} catch(std::bad_alloc&) {
		g_lastError = "memory allocation failed";
		return NEMO_ALLOCATION_ERROR;
	} catch(std::exception& e) {
		g_lastError = e.what();
		return NEMO_UNKNOWN_ERROR;
	} catch(...) {
		g_lastError = "unknown exception";
		return NEMO_UNKNOWN_ERROR;
	}
}

template<typename P>
void check(P* p, const char* what)
{
	if(!p) throw Error(NEMO_INVALID_INPUT, std::string("NULL ") + what + " handle");
}

} // namespace

extern "C" {

const char* nemo_version(void) { return "0.7.2-hip"; }

const char* nemo_strerror(void) { return g_lastError.c_str(); }

nemo_status_t nemo_add_plugin_path(const char* dir)
{
	return guarded([&] { nemo_hip::addPluginPath(dir); });
}

nemo_status_t nemo_hip_device_count(unsigned* count)
{
	return guarded([&] {
		int n = 0;
		if(hipGetDeviceCount(&n) != hipSuccess) n = 0;
		*count = unsigned(n);
	});
}

nemo_status_t nemo_hip_device_description(unsigned device, const char** descr)
{
	return guarded([&] {
		hipDeviceProp_t prop;
		HIP_CHECK(hipGetDeviceProperties(&prop, int(device)));
		g_descr = std::string(prop.name) + " (" + prop.gcnArchName + ", "
			+ std::to_string(prop.multiProcessorCount) + " CUs, "
			+ std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
		*descr = g_descr.c_str();
	});
}

/* ---------------------------------------------------------- configuration */

nemo_configuration_t nemo_new_configuration(void)
{
	nemo_configuration_t c = new nemo_configuration_class();
	/* default hardware (nemo.cpp:174-186): the HIP device if there is one */
	try {
		c->impl.setHipBackend(-1);
	} catch(...) {
		c->impl.description = "HIP backend (no device)";
	}
	return c;
}

void nemo_delete_configuration(nemo_configuration_t c) { delete c; }

nemo_status_t nemo_log_stdout(nemo_configuration_t c)
{
	return guarded([&] { check(c, "configuration"); c->impl.logging = true; });
}

nemo_status_t nemo_set_stdp_function(nemo_configuration_t c,
		float prefire_fn[], size_t prefire_len, float postfire_fn[], size_t postfire_len,
		float minE, float maxE, float minI, float maxI)
{
	return guarded([&] {
		check(c, "configuration");
		c->impl.setStdpFunction(std::vector<float>(prefire_fn, prefire_fn + prefire_len),
				std::vector<float>(postfire_fn, postfire_fn + postfire_len), minE, maxE, minI, maxI);
	});
}

nemo_status_t nemo_set_cpu_backend(nemo_configuration_t c)
{
	return guarded([&] {
		check(c, "configuration");
		c->impl.backend = NEMO_BACKEND_CPU;
		c->impl.description = "CPU backend (not provided by this library)";
	});
}

nemo_status_t nemo_set_hip_backend(nemo_configuration_t c, int dev)
{
	return guarded([&] { check(c, "configuration"); c->impl.setHipBackend(dev); });
}

nemo_status_t nemo_hip_device(nemo_configuration_t c, int* dev)
{
	return guarded([&] { check(c, "configuration"); *dev = c->impl.device; });
}

nemo_status_t nemo_backend(nemo_configuration_t c, backend_t* backend)
{
	return guarded([&] { check(c, "configuration"); *backend = c->impl.backend; });
}

nemo_status_t nemo_backend_description(nemo_configuration_t c, const char** descr)
{
	return guarded([&] { check(c, "configuration"); *descr = c->impl.description.c_str(); });
}

nemo_status_t nemo_set_write_only_synapses(nemo_configuration_t c)
{
	return guarded([&] { check(c, "configuration"); c->impl.writeOnlySynapses = true; });
}

/* ----------------------------------------------------------- construction */

nemo_network_t nemo_new_network(void) { return new nemo_network_class(); }

void nemo_delete_network(nemo_network_t n) { delete n; }

nemo_status_t nemo_add_neuron_type(nemo_network_t n, const char* name, unsigned* type)
{
	return guarded([&] { check(n, "network"); *type = n->impl.addNeuronType(name); });
}

nemo_status_t nemo_add_neuron_iz(nemo_network_t n, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma)
{
	/* Network::addNeuron (src/nemo/Network.cpp:47-56): args {a,b,c,d,sigma,u,v} */
	return guarded([&] {
		check(n, "network");
		unsigned type = n->impl.addNeuronType("Izhikevich");
		float args[7] = {a, b, c, d, sigma, u, v};
		n->impl.addNeurons(type, 1, &idx, 7, args);
	});
}

nemo_status_t nemo_add_neuron(nemo_network_t n, unsigned type, unsigned idx, unsigned nargs, float args[])
{
	return guarded([&] { check(n, "network"); n->impl.addNeurons(type, 1, &idx, nargs, args); });
}

nemo_status_t nemo_add_neurons(nemo_network_t n, unsigned type, size_t count,
		const unsigned idx[], unsigned nargs, const float args[])
{
	return guarded([&] { check(n, "network"); n->impl.addNeurons(type, count, idx, nargs, args); });
}

nemo_status_t nemo_add_synapse(nemo_network_t n, unsigned source, unsigned target,
		unsigned delay, float weight, unsigned char is_plastic, synapse_id* id)
{
	return guarded([&] {
		check(n, "network");
		n->impl.addSynapses(1, &source, &target, &delay, &weight, &is_plastic, id);
	});
}

nemo_status_t nemo_add_synapses(nemo_network_t n, size_t count,
		const unsigned source[], const unsigned target[], const unsigned delay[],
		const float weight[], const unsigned char plastic[], synapse_id ids[])
{
	return guarded([&] {
		check(n, "network");
		n->impl.addSynapses(count, source, target, delay, weight, plastic, ids);
	});
}

nemo_status_t nemo_neuron_count(nemo_network_t n, unsigned* ncount)
{
	return guarded([&] { check(n, "network"); *ncount = n->impl.neuronCount(); });
}

nemo_status_t nemo_get_neuron_state_n(nemo_network_t n, unsigned neuron, unsigned var, float* val)
{
	return guarded([&] { check(n, "network"); *val = n->impl.getNeuronState(neuron, var); });
}

nemo_status_t nemo_get_neuron_parameter_n(nemo_network_t n, unsigned neuron, unsigned param, float* val)
{
	return guarded([&] { check(n, "network"); *val = n->impl.getNeuronParameter(neuron, param); });
}

nemo_status_t nemo_set_neuron_n(nemo_network_t n, unsigned idx, unsigned nargs, float args[])
{
	return guarded([&] { check(n, "network"); n->impl.setNeuron(idx, nargs, args); });
}

/* Network::setNeuron (src/nemo/Network.cpp:67-73): args {a,b,c,d,sigma,u,v} */
nemo_status_t nemo_set_neuron_iz_n(nemo_network_t n, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma)
{
	return guarded([&] {
		check(n, "network");
		float args[7] = {a, b, c, d, sigma, u, v};
		n->impl.setNeuron(idx, 7, args);
	});
}

nemo_status_t nemo_set_neuron_state_n(nemo_network_t n, unsigned neuron, unsigned var, float val)
{
	return guarded([&] { check(n, "network"); n->impl.setNeuronState(neuron, var, val); });
}

nemo_status_t nemo_set_neuron_parameter_n(nemo_network_t n, unsigned neuron, unsigned param, float val)
{
	return guarded([&] { check(n, "network"); n->impl.setNeuronParameter(neuron, param, val); });
}

static size_t networkSynapse(nemo_network_t n, synapse_id id, const nemo_hip::Network::Axon** ax)
{
	*ax = &n->impl.axon(unsigned(id >> 32));
	size_t k = size_t(id & 0xffffffffu);
	if(k >= (*ax)->target.size()) throw Error(NEMO_INVALID_INPUT, "synapse id out or range");
	return k;
}

nemo_status_t nemo_get_synapse_source_n(nemo_network_t n, synapse_id id, unsigned* source)
{
	return guarded([&] {
		check(n, "network");
		const nemo_hip::Network::Axon* ax;
		networkSynapse(n, id, &ax);
		*source = unsigned(id >> 32);
	});
}

nemo_status_t nemo_get_synapse_target_n(nemo_network_t n, synapse_id id, unsigned* target)
{
	return guarded([&] {
		check(n, "network");
		const nemo_hip::Network::Axon* ax;
		size_t k = networkSynapse(n, id, &ax);
		*target = ax->target[k];
	});
}

nemo_status_t nemo_get_synapse_delay_n(nemo_network_t n, synapse_id id, unsigned* delay)
{
	return guarded([&] {
		check(n, "network");
		const nemo_hip::Network::Axon* ax;
		size_t k = networkSynapse(n, id, &ax);
		*delay = ax->delay[k];
	});
}

nemo_status_t nemo_get_synapse_weight_n(nemo_network_t n, synapse_id id, float* weight)
{
	return guarded([&] {
		check(n, "network");
		const nemo_hip::Network::Axon* ax;
		size_t k = networkSynapse(n, id, &ax);
		*weight = ax->weight[k];
	});
}

nemo_status_t nemo_get_synapse_plastic_n(nemo_network_t n, synapse_id id, unsigned char* plastic)
{
	return guarded([&] {
		check(n, "network");
		const nemo_hip::Network::Axon* ax;
		size_t k = networkSynapse(n, id, &ax);
		*plastic = ax->plastic[k];
	});
}

nemo_status_t nemo_get_synapses_from_n(nemo_network_t n, unsigned source, synapse_id* synapses[], size_t* len)
{
	return guarded([&] {
		check(n, "network");
		const std::vector<synapse_id>& ids = n->impl.getSynapsesFrom(source);
		*synapses = const_cast<synapse_id*>(ids.data());
		*len = ids.size();
	});
}

/* ------------------------------------------------------------- simulation */

nemo_simulation_t nemo_new_simulation(nemo_network_t n, nemo_configuration_t c)
{
	nemo_simulation_t s = nullptr;
	guarded([&] {
		check(n, "network");
		check(c, "configuration");
		s = new nemo_simulation_class(n->impl, c->impl);
	});
	return s;
}

void nemo_delete_simulation(nemo_simulation_t s) { delete s; }

nemo_status_t nemo_step(nemo_simulation_t s,
		unsigned fstim_nidx[], size_t fstim_count,
		unsigned istim_nidx[], float istim_current[], size_t istim_count,
		unsigned* fired[], size_t* fired_count)
{
	return guarded([&] {
		check(s, "simulation");
		const std::vector<unsigned>& f = s->impl.step(fstim_nidx, fstim_count, istim_nidx, istim_current, istim_count);
		if(fired) *fired = const_cast<unsigned*>(f.data());
		if(fired_count) *fired_count = f.size();
	});
}

nemo_status_t nemo_apply_stdp(nemo_simulation_t s, float reward)
{
	return guarded([&] { check(s, "simulation"); s->impl.applyStdp(reward); });
}

nemo_status_t nemo_get_membrane_potential(nemo_simulation_t s, unsigned neuron, float* v)
{
	return guarded([&] { check(s, "simulation"); *v = s->impl.getMembranePotential(neuron); });
}

nemo_status_t nemo_get_neuron_state_s(nemo_simulation_t s, unsigned neuron, unsigned var, float* val)
{
	return guarded([&] { check(s, "simulation"); *val = s->impl.getNeuronState(neuron, var); });
}

nemo_status_t nemo_get_neuron_parameter_s(nemo_simulation_t s, unsigned neuron, unsigned param, float* val)
{
	return guarded([&] { check(s, "simulation"); *val = s->impl.getNeuronParameter(neuron, param); });
}

nemo_status_t nemo_set_neuron_s(nemo_simulation_t s, unsigned idx, unsigned nargs, float args[])
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeuron(idx, nargs, args); });
}

nemo_status_t nemo_set_neurons_s(nemo_simulation_t s, size_t n, const unsigned idx[], unsigned nargs,
		const float args[])
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeurons(n, idx, nargs, args); });
}

nemo_status_t nemo_get_neuron_states_s(nemo_simulation_t s, size_t n, const unsigned idx[], unsigned var, float out[])
{
	return guarded([&] { check(s, "simulation"); s->impl.getNeuronValues(false, var, n, idx, out); });
}

nemo_status_t nemo_get_neuron_parameters_s(nemo_simulation_t s, size_t n, const unsigned idx[], unsigned param,
		float out[])
{
	return guarded([&] { check(s, "simulation"); s->impl.getNeuronValues(true, param, n, idx, out); });
}

nemo_status_t nemo_set_neuron_states_s(nemo_simulation_t s, size_t n, const unsigned idx[], unsigned var,
		const float val[])
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeuronValues(false, var, n, idx, val); });
}

nemo_status_t nemo_set_neuron_parameters_s(nemo_simulation_t s, size_t n, const unsigned idx[], unsigned param,
		const float val[])
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeuronValues(true, param, n, idx, val); });
}

/* Simulation::setNeuron (src/nemo/Simulation.cpp:21-27): args {a,b,c,d,sigma,u,v} */
nemo_status_t nemo_set_neuron_iz_s(nemo_simulation_t s, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma)
{
	return guarded([&] {
		check(s, "simulation");
		const float args[7] = {a, b, c, d, sigma, u, v};
		s->impl.setNeuron(idx, 7, args);
	});
}

nemo_status_t nemo_set_neuron_state_s(nemo_simulation_t s, unsigned neuron, unsigned var, float val)
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeuronState(neuron, var, val); });
}

nemo_status_t nemo_set_neuron_parameter_s(nemo_simulation_t s, unsigned neuron, unsigned param, float val)
{
	return guarded([&] { check(s, "simulation"); s->impl.setNeuronParameter(neuron, param, val); });
}

nemo_status_t nemo_get_synapse_source_s(nemo_simulation_t s, synapse_id id, unsigned* source)
{
	return guarded([&] {
		check(s, "simulation");
		(void) s->impl.getSynapseDelay(id); /* validates the id */
		*source = unsigned(id >> 32);
	});
}

nemo_status_t nemo_get_synapse_target_s(nemo_simulation_t s, synapse_id id, unsigned* target)
{
	return guarded([&] { check(s, "simulation"); *target = s->impl.getSynapseTarget(id); });
}

nemo_status_t nemo_get_synapse_delay_s(nemo_simulation_t s, synapse_id id, unsigned* delay)
{
	return guarded([&] { check(s, "simulation"); *delay = s->impl.getSynapseDelay(id); });
}

nemo_status_t nemo_get_synapse_weight_s(nemo_simulation_t s, synapse_id id, float* weight)
{
	return guarded([&] { check(s, "simulation"); *weight = s->impl.getSynapseWeight(id); });
}

nemo_status_t nemo_get_synapse_plastic_s(nemo_simulation_t s, synapse_id id, unsigned char* plastic)
{
	return guarded([&] { check(s, "simulation"); *plastic = s->impl.getSynapsePlastic(id); });
}

nemo_status_t nemo_get_synapses_from_s(nemo_simulation_t s, unsigned source, synapse_id* synapses[], size_t* len)
{
	return guarded([&] {
		check(s, "simulation");
		const std::vector<synapse_id>& ids = s->impl.getSynapsesFrom(source);
		*synapses = const_cast<synapse_id*>(ids.data());
		*len = ids.size();
	});
}

nemo_status_t nemo_elapsed_wallclock(nemo_simulation_t s, unsigned long* ms)
{
	return guarded([&] { check(s, "simulation"); *ms = s->impl.elapsedWallclock(); });
}

nemo_status_t nemo_elapsed_simulation(nemo_simulation_t s, unsigned long* ms)
{
	return guarded([&] { check(s, "simulation"); *ms = s->impl.elapsedSimulation(); });
}

nemo_status_t nemo_reset_timer(nemo_simulation_t s)
{
	return guarded([&] { check(s, "simulation"); s->impl.resetTimer(); });
}

nemo_status_t nemo_hip_run(nemo_simulation_t s, unsigned nsteps, unsigned stdp_period, float reward,
		unsigned counts[], unsigned* fired[], size_t* fired_len, uint64_t* events)
{
	return guarded([&] {
		check(s, "simulation");
		s->runCounts.clear();
		s->runFired.clear();
		const bool raster = counts != nullptr;
		nemo_hip::Simulation::FiredSink sink;
		sink.spill = &s->runFired;
		s->impl.run(nsteps, stdp_period, reward, raster, &s->runCounts, &sink, events);
		if(raster) {
			std::memcpy(counts, s->runCounts.data(), s->runCounts.size() * sizeof(unsigned));
			if(fired) *fired = s->runFired.data();
			if(fired_len) *fired_len = s->runFired.size();
		}
	});
}

nemo_status_t nemo_hip_run_into(nemo_simulation_t s, unsigned nsteps, unsigned stdp_period, float reward,
		unsigned counts[], unsigned fired_buf[], size_t fired_cap, size_t* fired_len, unsigned* spill[],
		size_t* spill_len, uint64_t* events)
{
	return guarded([&] {
		check(s, "simulation");
		if(!counts || !fired_len || !spill || !spill_len || (fired_cap && !fired_buf)) {
			throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		}
		s->runCounts.clear();
		s->runFired.clear();
		nemo_hip::Simulation::FiredSink sink;
		sink.buf = fired_buf;
		sink.cap = fired_cap;
		sink.spill = &s->runFired;
		s->impl.run(nsteps, stdp_period, reward, true, &s->runCounts, &sink, events);
		std::memcpy(counts, s->runCounts.data(), s->runCounts.size() * sizeof(unsigned));
		*fired_len = sink.len;
		*spill = s->runFired.data();
		*spill_len = s->runFired.size();
	});
}

nemo_status_t nemo_hip_new_batch(const nemo_network_t nets[], size_t count, nemo_configuration_t c,
		nemo_simulation_t* sim)
{
	return guarded([&] {
		check(c, "configuration");
		if(!sim || (!nets && count)) throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		std::vector<const nemo_hip::Network*> v(count);
		for(size_t k = 0; k < count; ++k) {
			check(nets[k], "network");
			v[k] = &nets[k]->impl;
		}
		*sim = new nemo_simulation_class(v, c->impl);
	});
}

nemo_status_t nemo_hip_batch_size(nemo_simulation_t s, unsigned* count)
{
	return guarded([&] { check(s, "simulation"); *count = s->impl.networkCount(); });
}

nemo_status_t nemo_hip_batch_get_neuron_state(nemo_simulation_t s, unsigned net, unsigned neuron, unsigned var,
		float* val)
{
	return guarded([&] { check(s, "simulation"); *val = s->impl.getNeuronState(neuron, var, net); });
}

nemo_status_t nemo_hip_batch_get_neuron_parameter(nemo_simulation_t s, unsigned net, unsigned neuron,
		unsigned param, float* val)
{
	return guarded([&] { check(s, "simulation"); *val = s->impl.getNeuronParameter(neuron, param, net); });
}

nemo_status_t nemo_hip_batch_get_membrane_potential(nemo_simulation_t s, unsigned net, unsigned neuron, float* val)
{
	return guarded([&] { check(s, "simulation"); *val = s->impl.getMembranePotential(neuron, net); });
}

nemo_status_t nemo_hip_prepare(nemo_simulation_t s, unsigned nsteps, unsigned stdp_period)
{
	return guarded([&] { check(s, "simulation"); s->impl.prepare(nsteps, stdp_period); });
}

static void shapeOut(const nemo_hip::LaunchShape& l, unsigned shape[NEMO_HIP_SHAPE_FIELDS])
{
	shape[0] = l.deliverBlocks;
	shape[1] = l.lateBlocks;
	shape[2] = l.binCapacity;
	shape[3] = l.stdpMode;
	shape[4] = l.stdpWords;
	shape[5] = l.overlap;
}

nemo_status_t nemo_hip_get_launch_shape(nemo_simulation_t s, unsigned shape[NEMO_HIP_SHAPE_FIELDS])
{
	return guarded([&] { check(s, "simulation"); shapeOut(s->impl.launchShape(), shape); });
}

nemo_status_t nemo_hip_set_launch_shape(nemo_simulation_t s, const unsigned shape[NEMO_HIP_SHAPE_FIELDS])
{
	return guarded([&] {
		check(s, "simulation");
		nemo_hip::LaunchShape l;
		l.deliverBlocks = shape[0];
		l.lateBlocks = shape[1];
		l.binCapacity = shape[2];
		l.stdpMode = shape[3];
		l.stdpWords = shape[4];
		l.overlap = shape[5];
		s->impl.setLaunchShape(l);
	});
}

nemo_status_t nemo_hip_autotune(nemo_simulation_t s, unsigned steps_per_trial, unsigned shape[NEMO_HIP_SHAPE_FIELDS])
{
	return guarded([&] {
		check(s, "simulation");
		const nemo_hip::LaunchShape l = s->impl.autotune(steps_per_trial);
		if(shape) shapeOut(l, shape);
	});
}

nemo_status_t nemo_hip_set_debug(nemo_simulation_t s, int enable)
{
	return guarded([&] { check(s, "simulation"); s->impl.setDebug(enable != 0); });
}

nemo_status_t nemo_hip_get_currents(nemo_simulation_t s, float E[], float I[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getCurrents(E, I, n); });
}

nemo_status_t nemo_hip_local_to_global(nemo_simulation_t s, unsigned gidx[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.localToGlobal(gidx, n); });
}

nemo_status_t nemo_hip_get_recent_firing(nemo_simulation_t s, uint64_t hist[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getRecentFiring(hist, n); });
}

nemo_status_t nemo_hip_get_rng_state(nemo_simulation_t s, uint32_t state[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getRngState(state, n); });
}

nemo_status_t nemo_hip_get_neuron_states(nemo_simulation_t s, unsigned var, float out[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getNeuronStates(var, out, n); });
}

nemo_status_t nemo_hip_get_fixed_weights(nemo_simulation_t s, int32_t w[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getFixedWeights(w, n); });
}

nemo_status_t nemo_hip_get_stdp_accumulators(nemo_simulation_t s, int32_t acc[], size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.getStdpAccumulators(acc, n); });
}

nemo_status_t nemo_hip_counts(nemo_simulation_t s, size_t* neurons, size_t* synapses, size_t* plastic)
{
	return guarded([&] {
		check(s, "simulation");
		if(neurons) *neurons = s->impl.neuronCount();
		if(synapses) *synapses = s->impl.synapseCount();
		if(plastic) *plastic = s->impl.plasticCount();
	});
}

nemo_status_t nemo_hip_profile(nemo_simulation_t s, unsigned nsteps, double avg_us[4], uint64_t counters[4])
{
	return guarded([&] { check(s, "simulation"); s->impl.profile(nsteps, avg_us, counters); });
}

nemo_status_t nemo_hip_eval_math(int fn, const float in[], float out[], size_t n)
{
	return guarded([&] {
		if(fn < 0 || fn > 2) throw Error(NEMO_INVALID_INPUT, "unknown function");
		if(n == 0) return;
		nemo_hip::DeviceArray<float> din(n), dout(n);
		HIP_CHECK(hipMemcpy(din.ptr, in, n * sizeof(float), hipMemcpyHostToDevice));
		HIP_CHECK(nemo_launch_eval_math(nullptr, fn, din.ptr, dout.ptr, n));
		HIP_CHECK(hipDeviceSynchronize());
		HIP_CHECK(hipMemcpy(out, dout.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
	});
}

nemo_status_t nemo_hip_propagate(nemo_simulation_t s, const unsigned fired[], size_t nfired, float E[], float I[],
		size_t n)
{
	return guarded([&] { check(s, "simulation"); s->impl.propagate(fired, nfired, E, I, n); });
}

nemo_status_t nemo_hip_propagate_device(nemo_simulation_t s, const uint32_t* fired, float** E, float** I)
{
	return guarded([&] { check(s, "simulation"); s->impl.propagateDevice(fired, E, I); });
}

/* cuda_test_device (src/nemo/cuda/devices.hpp:33-37; devices.cpp), as a
 * status: the device exists and accepts an allocation */
nemo_status_t nemo_hip_test_device(unsigned device)
{
	return guarded([&] {
		int n = 0;
		HIP_CHECK(hipGetDeviceCount(&n));
		if(int(device) >= n) throw Error(NEMO_INVALID_INPUT, "invalid HIP device");
		HIP_CHECK(hipSetDevice(int(device)));
		void* p = nullptr;
		HIP_CHECK(hipMalloc(&p, 256));
		HIP_CHECK(hipFree(p));
	});
}

} // extern "C"

/* ---------------------------------------------------- float-weight mode */

struct nemo_hip_float_network_class {
	nemo_hip_float_network_class(const nemo_hip::Network& n, int dev) : impl(n, dev) {}
	nemo_hip::FloatNetwork impl;
};

extern "C" {

nemo_status_t nemo_hip_new_float_network(nemo_network_t n, int device, nemo_hip_float_network_t* out)
{
	return guarded([&] {
		check(n, "network");
		if(!out) throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		*out = new nemo_hip_float_network_class(n->impl, device);
	});
}

void nemo_hip_delete_float_network(nemo_hip_float_network_t f)
{
	delete f;
}

nemo_status_t nemo_hip_float_network_sizes(nemo_hip_float_network_t f, size_t* neurons, size_t* synapses,
		unsigned* max_delay)
{
	return guarded([&] {
		check(f, "float network");
		if(neurons) *neurons = f->impl.neurons();
		if(synapses) *synapses = f->impl.synapses();
		if(max_delay) *max_delay = f->impl.maxDelay();
	});
}

nemo_status_t nemo_hip_float_forward(nemo_hip_float_network_t f, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* v0, const float* u0, const float* x,
		float* spikes, float* v_out, double* checkpoints)
{
	return guarded([&] {
		check(f, "float network");
		if(!w || !a || !b || !c || !d || !v0 || !u0 || !x || !spikes || !v_out || !checkpoints) {
			throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		}
		f->impl.forward(static_cast<hipStream_t>(stream), steps, checkpoint_every, w, a, b, c, d, v0, u0, x,
				spikes, v_out, checkpoints);
	});
}

nemo_status_t nemo_hip_float_forward_static(nemo_hip_float_network_t f, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* v0, const float* u0, const float* x,
		float* spikes, float* v_out, double* checkpoints)
{
	return guarded([&] {
		check(f, "float network");
		if(!w || !a || !b || !c || !d || !v0 || !u0 || !x || !spikes || !v_out || !checkpoints) {
			throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		}
		f->impl.forwardStatic(static_cast<hipStream_t>(stream), steps, checkpoint_every, w, a, b, c, d, v0, u0, x,
				spikes, v_out, checkpoints);
	});
}

nemo_status_t nemo_hip_float_backward(nemo_hip_float_network_t f, void* stream, unsigned steps,
		unsigned checkpoint_every, const float* w, const float* a, const float* b, const float* c,
		const float* d, const float* x, const float* spikes, const double* checkpoints,
		const float* grad_spikes, const float* grad_v, float* grad_w, float* grad_a, float* grad_b,
		float* grad_d, float* grad_x, float* grad_v0, float* grad_u0)
{
	return guarded([&] {
		check(f, "float network");
		if(!w || !a || !b || !c || !d || !x || !spikes || !checkpoints || !grad_w || !grad_a || !grad_b
				|| !grad_d) {
			throw nemo_hip::Error(NEMO_INVALID_INPUT, "null argument");
		}
		f->impl.backward(static_cast<hipStream_t>(stream), steps, checkpoint_every, w, a, b, c, d, x, spikes,
				checkpoints, grad_spikes, grad_v, grad_w, grad_a, grad_b, grad_d, grad_x, grad_v0, grad_u0);
	});
}

} // extern "C"
