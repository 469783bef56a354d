/*
 * CPU oracle: a Boost-free restatement of NeMo's CPU simulation backend.
 *
 * TEST INFRASTRUCTURE ONLY -- see nemo_oracle.h.  Every function below cites
 * the reference file and lines whose behaviour it restates.  The reference is
 * NeMo v0.7.2; paths are relative to its source tree.
 *
 * What is restated, and where:
 *   - fixed point Q(31-f).f            src/nemo/fixedpoint.cpp:17-49, fixedpoint.hpp:37-50
 *   - xorshift128 urand / Box-Muller   src/nemo/RNG.cpp:9-33
 *   - per-neuron seeding (mt19937)     src/nemo/RNG.cpp:38-63
 *   - network storage and ordering     src/nemo/NetworkImpl.cpp:27-128, Neurons.cpp:20-40, Axon.cpp:9-20
 *   - local index mapping              src/nemo/cpu/Simulation.cpp:60-77, cpu/Neurons.cpp:7-48
 *   - forward rows (source, delay)     src/nemo/ConnectivityMatrix.cpp:65-176
 *   - reverse (plastic) matrix         src/nemo/construction/RCM.cpp:112-178, runtime/RCM.cpp
 *   - STDP function and process        src/nemo/StdpFunction.cpp:12-108, StdpProcess.cpp:20-138
 *   - STDP accumulate / apply          src/nemo/ConnectivityMatrix.cpp:224-325
 *   - spike delivery and currents      src/nemo/cpu/Simulation.cpp:259-308
 *   - step protocol                    src/nemo/SimulationBackend.cpp:22-77, cpu/Simulation.cpp:94-111
 *   - neuron models                    src/nemo/cpu/plugins/{Izhikevich,QIF,IF_curr_exp,HH,PoissonSource,Input}.cpp
 *
 * Deliberate deviations:
 *   1. QIF and HH index their per-type parameter/state arrays with the local
 *      index nl = n - start.  The reference indexes them with the simulation
 *      index n (QIF.cpp:53-105, HH.cpp:68-135), which reads out of bounds
 *      unless the type starts at local index 0; for start == 0 both agree.
 *   2. A network maximum delay of 64 delivers spikes of every delay.  The
 *      reference computes ~(~0 << 64) (cpu/Simulation.cpp:264), which is
 *      undefined behaviour in C++.
 *   3. Signed 32-bit overflow in STDP accumulation and weight update wraps
 *      (two's complement), which is what the reference's x86 build does.
 *   4. In STDP accumulation a presynaptic history shifted by a delay of 64
 *      is empty.  The reference shifts a 64-bit word by the delay
 *      (ConnectivityMatrix.cpp:248), undefined behaviour for 64; the
 *      case needs a postfire window of length 0.
 *
 * Build: g++ -O2 -fopenmp -ffp-contract=off (no fast-math), see Makefile.
 */

#include "nemo_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

/* status codes, src/nemo/types.h:31-61 */
#define NEMO_OK 0
#define NEMO_API_UNSUPPORTED 5
#define NEMO_INVALID_INPUT 6
#define NEMO_BUFFER_UNDERFLOW 8
#define NEMO_LOGIC_ERROR 9
#define NEMO_UNKNOWN_ERROR 14

namespace {

typedef int32_t fix_t;
typedef int64_t wfix_t;

struct oracle_error : std::runtime_error {
	int code;
	oracle_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

std::string g_lastError;

/* ------------------------------------------------------------------ */
/* fixed point: src/nemo/fixedpoint.cpp:17-49, fixedpoint.hpp:37-50   */
/* ------------------------------------------------------------------ */

const fix_t fx_min = INT32_MIN;
const fix_t fx_max = INT32_MAX;

fix_t fx_toFix(float f, unsigned fbits)
{
	/* fixedpoint.cpp:20: overflow when |int(f)| >= 2^(31-fbits) */
	if(std::abs(int(f)) >= (1 << (31 - fbits))) {
		throw oracle_error(NEMO_INVALID_INPUT, "Fixed-point overflow");
	}
	return static_cast<fix_t>(f * float(1 << fbits));
}

float fx_toFloat(fix_t v, unsigned fbits)
{
	return float(v) / float(1 << fbits);
}

float wfx_toFloat(wfix_t v, unsigned fbits)
{
	/* NEMO_WEIGHT_FIXED_POINT_SATURATION is on by default */
	v = std::min(v, wfix_t(fx_max));
	v = std::max(v, wfix_t(fx_min));
	return float(v) / float(1 << fbits);
}

fix_t fx_mul(fix_t a, fix_t b, unsigned fbits)
{
	int64_t r = int64_t(a) * int64_t(b);
	return fix_t(r >> fbits);
}

fix_t wrap_add(fix_t a, fix_t b)
{
	return fix_t(uint32_t(a) + uint32_t(b));
}

/* ------------------------------------------------------------------ */
/* RNG: src/nemo/RNG.cpp:9-63                                          */
/* ------------------------------------------------------------------ */

struct RNG { uint32_t state[4]; };

uint32_t urand(RNG* rng)
{
	/* Marsaglia xorshift128, RNG.cpp:10-18 */
	uint32_t t = (rng->state[0] ^ (rng->state[0] << 11));
	rng->state[0] = rng->state[1];
	rng->state[1] = rng->state[2];
	rng->state[2] = rng->state[3];
	rng->state[3] = (rng->state[3] ^ (rng->state[3] >> 19)) ^ (t ^ (t >> 8));
	return rng->state[3];
}

float nrand(RNG* rng)
{
	/* Box-Muller, sine branch only, RNG.cpp:25-33.  Host glibc logf/sinf/sqrtf. */
	float a = urand(rng) * 1.4629180792671596810513378043098e-9f;
	float b = urand(rng) * 0.00000000023283064365386962890625f;
	float r = sqrtf(-2 * logf(1 - b));
	return sinf(a) * r;
}

/* RNG.cpp:38-63.  boost::mt19937 default-seeded (5489) feeds
 * boost::uniform_int<unsigned long>(0, 0x7fffffff); for a 32-bit engine and
 * this range Boost >= 1.47 uses bucket size 2, i.e. each seed word is
 * mt() >> 1.  Four words per neuron, the first 4*minIdx words skipped. */
void initialiseRng(unsigned minIdx, unsigned maxIdx, std::vector<RNG>& rngs)
{
	if(minIdx > maxIdx) {
		throw oracle_error(NEMO_LOGIC_ERROR, "Invalid neuron range when initialising RNG");
	}
	std::mt19937 mt; /* default seed 5489, same algorithm as boost::mt19937 */
	for(unsigned i = 0; i < 4 * minIdx; ++i) {
		mt();
	}
	for(unsigned g = minIdx; g <= maxIdx; ++g) {
		RNG& s = rngs.at(g - minIdx);
		for(unsigned plane = 0; plane < 4; ++plane) {
			s.state[plane] = uint32_t(mt()) >> 1;
		}
	}
}

/* ------------------------------------------------------------------ */
/* neuron types: the .ini files in src/nemo/plugins/                   */
/* ------------------------------------------------------------------ */

enum Model { IZHIKEVICH, QIF, IF_CURR_EXP, HH, POISSON, INPUT, KURAMOTO, KURAMOTO_LAG, IZHIKEVICH_CHAOS,
	IZHIKEVICH_SYN, IZHIKEVICH_2007, IZHIKEVICH_2007_DOPSYN, BAYESIAN, INHOMOGENEOUS_POISSON };

struct NeuronType {
	std::string name;
	Model model;
	unsigned nParam;
	unsigned nState;
	unsigned membranePotential;
	unsigned history = 1; /* state history length (NeuronType.hpp:67) */
	bool rcm = false;     /* reads sources, delays and weights of all incoming synapses */
};

NeuronType lookupType(const std::string& name)
{
	/* parameter/state counts and membrane-potential index from the .ini files:
	 * Izhikevich.ini:2-4, QIF.ini:2-4, IF_curr_exp.ini:14-29, HH.ini:2-4,
	 * PoissonSource.ini:1-3, Input.ini:1-3 */
	if(name == "Izhikevich")    return NeuronType{name, IZHIKEVICH, 5, 2, 1};
	if(name == "QIF")           return NeuronType{name, QIF, 1, 1, 0};
	if(name == "IF_curr_exp")   return NeuronType{name, IF_CURR_EXP, 9, 4, 0};
	if(name == "HH")            return NeuronType{name, HH, 0, 5, 0};
	if(name == "PoissonSource") return NeuronType{name, POISSON, 1, 0, 0};
	if(name == "Input")         return NeuronType{name, INPUT, 0, 0, 0};
	/* Kuramoto.ini:2-20: history=32, rcm sources/delays/weights */
	if(name == "Kuramoto")      return NeuronType{name, KURAMOTO, 2, 1, 0, 32, true};
	/* KuramotoLag.ini:1-14: a third parameter, the phase lag alpha */
	if(name == "KuramotoLag")   return NeuronType{name, KURAMOTO_LAG, 3, 1, 0, 32, true};
	/* the CPU-only models: IzhikevichChaos.ini, IzhikevichSyn.h (12 parameters),
	 * Izhikevich2007.ini, Izhikevich2007DopSyn.ini, BayesianNeuron.ini,
	 * InhomogeneousPoissonSource.ini */
	if(name == "IzhikevichChaos")      return NeuronType{name, IZHIKEVICH_CHAOS, 5, 2, 1};
	if(name == "IzhikevichSyn")        return NeuronType{name, IZHIKEVICH_SYN, 12, 4, 1};
	if(name == "Izhikevich2007")       return NeuronType{name, IZHIKEVICH_2007, 16, 4, 1};
	if(name == "Izhikevich2007DopSyn") return NeuronType{name, IZHIKEVICH_2007_DOPSYN, 18, 4, 1};
	if(name == "BayesianNeuron")       return NeuronType{name, BAYESIAN, 3, 2, 1};
	if(name == "InhomogeneousPoissonSource") return NeuronType{name, INHOMOGENEOUS_POISSON, 0, 0, 0};
	throw oracle_error(NEMO_INVALID_INPUT, "unknown neuron type " + name);
}

} // namespace

/* ------------------------------------------------------------------ */
/* network: src/nemo/NetworkImpl.cpp, Neurons.cpp, Axon.cpp            */
/* ------------------------------------------------------------------ */

struct oracle_network {
	std::vector<NeuronType> types;
	std::map<std::string, unsigned> typeIds;
	/* per type: param[p][i], state[s][i], gidx[i] in insertion order (Neurons.cpp:20-40) */
	struct Group {
		std::vector<std::vector<float> > param;
		std::vector<std::vector<float> > state;
		std::vector<unsigned> gidx;
	};
	std::vector<Group> groups;
	std::map<unsigned, std::pair<unsigned, unsigned> > mapper; /* gidx -> (type, index) */

	/* synapses grouped by source in insertion order (NetworkImpl.hpp:154, Axon.cpp:9-20) */
	struct Axon {
		std::vector<unsigned> target, delay;
		std::vector<float> weight;
		std::vector<unsigned char> plastic;
	};
	std::map<unsigned, Axon> fcm;
	unsigned maxDelay = 0;
	size_t synapseCount = 0;
};

struct oracle_configuration {
	bool stdp = false;
	std::vector<float> prefire, postfire;
	float minE = 0, maxE = 0, minI = 0, maxI = 0;
	unsigned fbits = 20; /* ConfigurationImpl.cpp:22 */
};

namespace {

/* StdpProcess, src/nemo/StdpProcess.cpp:20-138, with the bit masks of
 * StdpFunction::getBits (StdpFunction.cpp:72-108) */
struct StdpProcess {
	std::vector<fix_t> fnPre, fnPost;
	int postPreWindow = 0;
	uint64_t potentiationBits = 0, depressionBits = 0;
	uint64_t prePostBits = 0, postPreBits = 0, postFireMask = 0;
	fix_t minE = 0, maxE = 0, minI = 0, maxI = 0;
	static const unsigned NO_APPLICATION = ~0U;

	static uint64_t lowMask(unsigned n) { return n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1); }

	StdpProcess() {}
	StdpProcess(const oracle_configuration& c) {
		unsigned fbits = c.fbits;
		minE = fx_toFix(c.minE, fbits);
		maxE = fx_toFix(c.maxE, fbits);
		minI = fx_toFix(c.minI, fbits);
		maxI = fx_toFix(c.maxI, fbits);
		for(float f : c.prefire) fnPre.push_back(fx_toFix(f, fbits));
		for(float f : c.postfire) fnPost.push_back(fx_toFix(f, fbits));
		/* StdpFunction::getBits: postfire reversed first (bits 0..), then prefire */
		unsigned n = 0;
		for(auto f = c.postfire.rbegin(); f != c.postfire.rend(); ++f, ++n) {
			if(*f > 0.0f) potentiationBits |= uint64_t(1) << n;
			if(*f < 0.0f) depressionBits |= uint64_t(1) << n;
		}
		for(auto f = c.prefire.begin(); f != c.prefire.end(); ++f, ++n) {
			if(*f > 0.0f) potentiationBits |= uint64_t(1) << n;
			if(*f < 0.0f) depressionBits |= uint64_t(1) << n;
		}
		int prePostWindow = int(c.prefire.size());
		postPreWindow = int(c.postfire.size());
		prePostBits = lowMask(prePostWindow) << postPreWindow;
		postPreBits = lowMask(postPreWindow);
		postFireMask = uint64_t(1) << postPreWindow;
	}

	unsigned closestPreFire(uint64_t arrivals) const {
		uint64_t valid = arrivals & prePostBits;
		return valid ? unsigned(__builtin_ctzll(valid >> postPreWindow)) : NO_APPLICATION;
	}

	unsigned closestPostFire(uint64_t arrivals) const {
		uint64_t valid = arrivals & postPreBits;
		return valid ? unsigned(__builtin_clzll(valid << (64 - postPreWindow))) : NO_APPLICATION;
	}

	fix_t updateRegion(uint64_t arrivals) const {
		fix_t w_diff = 0;
		if(arrivals) {
			unsigned dt_pre = closestPreFire(arrivals);
			unsigned dt_post = closestPostFire(arrivals);
			if(dt_pre < dt_post) {
				w_diff = fnPre.at(dt_pre);
			} else if(dt_post < dt_pre) {
				w_diff = fnPost.at(dt_post);
			}
		}
		return w_diff;
	}

	fix_t weightChange(uint64_t preFiring) const {
		return wrap_add(updateRegion(preFiring & potentiationBits),
				updateRegion(preFiring & depressionBits));
	}

	fix_t updatedWeight(fix_t w_old, fix_t w_diff) const {
		fix_t w_new = 0;
		if(w_old > 0) {
			w_new = std::min(maxE, std::max(wrap_add(w_old, w_diff), minE));
		} else if(w_old < 0) {
			w_new = std::min(minI, std::max(wrap_add(w_old, w_diff), maxI));
		}
		return w_new;
	}
};

struct FAxonTerminal { uint32_t target; fix_t weight; };

} // namespace

/* ------------------------------------------------------------------ */
/* simulation: src/nemo/cpu/Simulation.cpp                             */
/* ------------------------------------------------------------------ */

struct oracle_simulation {
	unsigned fbits = 20;
	size_t ncount = 0;

	/* mapper: src/nemo/RandomMapper.hpp, filled as in cpu/Neurons.cpp:27-40 */
	std::unordered_map<unsigned, unsigned> g2l;
	std::vector<unsigned> l2g;

	struct TypeGroup {
		NeuronType type;
		unsigned base, size;
		std::vector<float> param; /* [nParam][size] */
		std::vector<float> state; /* [history][nState][size] (cpu/Neurons.cpp:15) */
		unsigned current = 0;     /* history slot of the current state (cpu/Neurons.cpp:123) */
		std::vector<RNG> rng;
		float& at(unsigned var, unsigned i) { return state[(size_t(current) * type.nState + var) * size + i]; }
	};
	std::vector<TypeGroup> groups;
	std::vector<unsigned> typeOfLocal;

	std::vector<unsigned> fired;
	std::vector<uint64_t> recentFiring;
	std::vector<uint64_t> delays;
	std::vector<wfix_t> fxE, fxI;
	std::vector<float> currentE, currentI, currentExt;
	std::vector<unsigned> fstim;

	/* forward matrix: rows addressed by source*maxDelay + delay-1
	 * (ConnectivityMatrix.cpp:147-176, ConnectivityMatrix.hpp:288-292) */
	unsigned maxDelay = 0;
	std::vector<size_t> rowStart; /* size nrows+1 */
	std::vector<FAxonTerminal> fcm;

	/* auxiliary per-synapse data for queries, keyed by global source
	 * (ConnectivityMatrix.cpp:109-117): fcm address of each synapse */
	std::map<unsigned, std::pair<size_t, size_t> > auxRange; /* gsrc -> [first, last) in synapse order */
	std::vector<size_t> auxAddr;     /* synapse order -> fcm index */
	std::vector<unsigned> auxDelay;
	std::vector<unsigned char> auxPlastic;

	/* reverse matrix for STDP: plastic synapses grouped by target
	 * (construction/RCM.cpp:112-178) */
	bool stdpEnabled = false;
	StdpProcess stdp;
	std::vector<size_t> rcmStart; /* per local target, size ncount+1 */
	std::vector<uint32_t> rcmSource, rcmDelay;
	std::vector<size_t> rcmForward; /* fcm index of the synapse's weight */
	std::vector<fix_t> rcmAcc;

	/* reverse matrix of all synapses, for models that read presynaptic
	 * state (construction/RCM.cpp:60-98 with rcm sources/delays/weights):
	 * per local target, source, delay and float weight in the network's
	 * synapse order */
	std::vector<size_t> rcmAllStart;
	std::vector<uint32_t> rcmAllSource, rcmAllDelay;
	std::vector<float> rcmAllWeight;

	std::vector<unsigned> firedOut;
	unsigned long cycle = 0; /* Timer::elapsedSimulation, reset by resetTimer */
	uint64_t delivered = 0;

	unsigned localIdx(unsigned g) const {
		auto i = g2l.find(g);
		if(i == g2l.end()) {
			throw oracle_error(NEMO_INVALID_INPUT, "Non-existing neuron index " + std::to_string(g));
		}
		return i->second;
	}
};

namespace {

void updateIzhikevich(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateQIF(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateIFCurrExp(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateHH(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updatePoisson(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateInput(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateKuramoto(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateIzhikevichVariant(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateBayesian(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void updateInhomogeneousPoisson(oracle_simulation& s, oracle_simulation::TypeGroup& g);
void initKuramoto(oracle_simulation::TypeGroup& g);

void build(oracle_simulation& s, const oracle_network& net, const oracle_configuration& conf)
{
	/* nemo.cpp:111-117: empty network and STDP window checks */
	size_t ncount = 0;
	for(auto& g : net.groups) ncount += g.gidx.size();
	if(ncount == 0) {
		throw oracle_error(NEMO_INVALID_INPUT, "Cannot create simulation from empty network");
	}
	if(conf.stdp && net.maxDelay + conf.postfire.size() > 64) {
		throw oracle_error(NEMO_INVALID_INPUT, "Max delay + postfire part of STDP function must be <= 64");
	}
	/* cpu/Simulation.cpp:52-56 */
	if(net.maxDelay > 64) {
		throw oracle_error(NEMO_INVALID_INPUT, "The CPU backend supports a maximum of 64 ms");
	}

	s.fbits = conf.fbits;
	s.ncount = ncount;

	/* cpu/Simulation.cpp:60-77 and cpu/Neurons.cpp:7-48: contiguous local
	 * indices, type by type, insertion order within a type */
	unsigned l_idx = 0;
	s.l2g.resize(ncount);
	s.typeOfLocal.resize(ncount);
	for(unsigned t = 0; t < net.types.size(); ++t) {
		const auto& ng = net.groups[t];
		if(ng.gidx.empty()) continue;
		oracle_simulation::TypeGroup tg;
		tg.type = net.types[t];
		tg.base = l_idx;
		tg.size = unsigned(ng.gidx.size());
		tg.param.assign(size_t(tg.type.nParam) * tg.size, 0.0f);
		tg.state.assign(size_t(tg.type.history) * tg.type.nState * tg.size, 0.0f);
		for(unsigned i = 0; i < tg.size; ++i) {
			unsigned sim = tg.base + i;
			if(!s.g2l.emplace(ng.gidx[i], sim).second) {
				throw oracle_error(NEMO_INVALID_INPUT, "Duplicate neuron index " + std::to_string(ng.gidx[i]));
			}
			s.l2g[sim] = ng.gidx[i];
			s.typeOfLocal[sim] = unsigned(s.groups.size());
			for(unsigned p = 0; p < tg.type.nParam; ++p) tg.param[size_t(p) * tg.size + i] = ng.param[p][i];
			for(unsigned v = 0; v < tg.type.nState; ++v) tg.state[size_t(v) * tg.size + i] = ng.state[v][i];
		}
		tg.rng.resize(tg.size);
		initialiseRng(tg.base, tg.base + tg.size - 1, tg.rng);
		if(tg.type.model == KURAMOTO || tg.type.model == KURAMOTO_LAG) initKuramoto(tg);
		l_idx += tg.size;
		s.groups.push_back(std::move(tg));
	}

	s.fired.assign(ncount, 0);
	s.recentFiring.assign(ncount, 0);
	s.delays.assign(ncount, 0);
	s.fxE.assign(ncount, 0);
	s.fxI.assign(ncount, 0);
	s.currentE.assign(ncount, 0.0f);
	s.currentI.assign(ncount, 0.0f);
	s.currentExt.assign(ncount, 0.0f);
	s.fstim.assign(ncount, 0);

	/* ConnectivityMatrix.cpp:65-176.  Synapses are visited in the network's
	 * iteration order: source ascending, then insertion order
	 * (NetworkImpl.cpp:205-236).  Rows (local source, delay) keep that order. */
	s.stdpEnabled = conf.stdp;
	if(conf.stdp) {
		s.stdp = StdpProcess(conf);
	}
	s.maxDelay = net.maxDelay;
	const unsigned D = std::max(1u, net.maxDelay);
	size_t nrows = ncount * D;
	std::vector<size_t> rowCount(nrows + 1, 0);
	std::vector<unsigned> rcmCount(ncount + 1, 0);
	/* construction/RCM.cpp:83-91: STDP and a model that reads the reverse
	 * matrix do not mix */
	bool rcmAll = false;
	for(const auto& g : s.groups) rcmAll = rcmAll || g.type.rcm;
	if(rcmAll && conf.stdp) {
		throw oracle_error(NEMO_API_UNSUPPORTED, "The current version does not support a mix of STDP with neuron "
				"types which require reverse connectivity for other purposes");
	}
	std::vector<unsigned> rcmAllCount(ncount + 1, 0);

	/* pass 1: validate, count rows */
	for(auto& kv : net.fcm) {
		unsigned src = s.localIdx(kv.first);
		const auto& ax = kv.second;
		for(size_t k = 0; k < ax.target.size(); ++k) {
			unsigned tgt = s.localIdx(ax.target[k]);
			rowCount[size_t(src) * D + ax.delay[k] - 1] += 1;
			if(conf.stdp && ax.plastic[k]) rcmCount[tgt] += 1;
			if(rcmAll) rcmAllCount[tgt] += 1;
		}
	}
	if(rcmAll) {
		s.rcmAllStart.assign(ncount + 1, 0);
		for(size_t n = 0; n < ncount; ++n) s.rcmAllStart[n + 1] = s.rcmAllStart[n] + rcmAllCount[n];
		s.rcmAllSource.resize(s.rcmAllStart[ncount]);
		s.rcmAllDelay.resize(s.rcmAllStart[ncount]);
		s.rcmAllWeight.resize(s.rcmAllStart[ncount]);
	}
	std::vector<size_t> rcmAllFill(s.rcmAllStart.begin(),
			s.rcmAllStart.empty() ? s.rcmAllStart.begin() : s.rcmAllStart.end() - 1);
	s.rowStart.assign(nrows + 1, 0);
	for(size_t r = 0; r < nrows; ++r) s.rowStart[r + 1] = s.rowStart[r] + rowCount[r];
	s.fcm.resize(net.synapseCount);
	std::vector<size_t> rowFill(s.rowStart.begin(), s.rowStart.end() - 1);

	if(conf.stdp) {
		s.rcmStart.assign(ncount + 1, 0);
		for(size_t n = 0; n < ncount; ++n) s.rcmStart[n + 1] = s.rcmStart[n] + rcmCount[n];
		size_t np = s.rcmStart[ncount];
		s.rcmSource.resize(np);
		s.rcmDelay.resize(np);
		s.rcmForward.resize(np);
		s.rcmAcc.assign(np, 0);
	}
	std::vector<size_t> rcmFill(s.rcmStart.begin(), s.rcmStart.empty() ? s.rcmStart.begin() : s.rcmStart.end() - 1);

	s.auxAddr.resize(net.synapseCount);
	s.auxDelay.resize(net.synapseCount);
	s.auxPlastic.resize(net.synapseCount);
	size_t seq = 0;
	/* pass 2: fill rows, reverse matrix and the auxiliary query map */
	for(auto& kv : net.fcm) {
		unsigned src = s.localIdx(kv.first);
		const auto& ax = kv.second;
		s.auxRange[kv.first] = std::make_pair(seq, seq + ax.target.size());
		for(size_t k = 0; k < ax.target.size(); ++k, ++seq) {
			unsigned tgt = s.localIdx(ax.target[k]);
			unsigned d = ax.delay[k];
			fix_t w = fx_toFix(ax.weight[k], s.fbits);
			size_t addr = rowFill[size_t(src) * D + d - 1]++;
			s.fcm[addr] = FAxonTerminal{tgt, w};
			s.delays[src] |= uint64_t(1) << (d - 1); /* OutgoingDelays.cpp:74-84 */
			s.auxAddr[seq] = addr;
			s.auxDelay[seq] = d;
			s.auxPlastic[seq] = ax.plastic[k] != 0;
			if(conf.stdp && ax.plastic[k]) {
				size_t r = rcmFill[tgt]++;
				s.rcmSource[r] = src;
				s.rcmDelay[r] = d;
				s.rcmForward[r] = addr;
			}
			if(rcmAll) {
				size_t r = rcmAllFill[tgt]++;
				s.rcmAllSource[r] = src;
				s.rcmAllDelay[r] = d;
				s.rcmAllWeight[r] = ax.weight[k];
			}
		}
	}
	s.cycle = 0;
}

/* cpu/Simulation.cpp:259-308 */
void deliverSpikes(oracle_simulation& s)
{
	uint64_t validSpikes = s.maxDelay >= 64 ? ~uint64_t(0) : ~((~uint64_t(0)) << s.maxDelay);
	const unsigned D = std::max(1u, s.maxDelay);
	for(size_t source = 0; source < s.ncount; ++source) {
		uint64_t f = s.recentFiring[source] & validSpikes & s.delays[source];
		int delay = 0;
		while(f) {
			int shift = 1 + __builtin_ctzll(f);
			delay += shift;
			f = shift >= 64 ? 0 : (f >> shift);
			/* deliverSpikesOne, cpu/Simulation.cpp:293-308 */
			size_t r = source * D + unsigned(delay) - 1;
			for(size_t k = s.rowStart[r]; k < s.rowStart[r + 1]; ++k) {
				const FAxonTerminal& term = s.fcm[k];
				std::vector<wfix_t>& current = term.weight >= 0 ? s.fxE : s.fxI;
				current[term.target] += term.weight;
			}
			s.delivered += s.rowStart[r + 1] - s.rowStart[r];
		}
	}
	/* cpu/Simulation.cpp:279-288 */
	const unsigned fbits = s.fbits;
	const int ncount = int(s.ncount);
#pragma omp parallel for default(shared)
	for(int n = 0; n < ncount; n++) {
		s.currentE[n] = wfx_toFloat(s.fxE[n], fbits);
		s.fxE[n] = 0;
		s.currentI[n] = wfx_toFloat(s.fxI[n], fbits);
		s.fxI[n] = 0;
	}
}

/* ConnectivityMatrix.cpp:224-258 */
void accumulateStdp(oracle_simulation& s)
{
	if(!s.stdpEnabled) return;
	for(size_t target = 0; target < s.ncount; ++target) {
		if(!(s.recentFiring[target] & s.stdp.postFireMask)) continue;
		for(size_t r = s.rcmStart[target]; r < s.rcmStart[target + 1]; ++r) {
			const unsigned dr = s.rcmDelay[r];
			uint64_t preFiring = dr >= 64 ? 0 : s.recentFiring[s.rcmSource[r]] >> dr;
			fix_t w_diff = s.stdp.weightChange(preFiring);
			if(w_diff != 0) {
				s.rcmAcc[r] = wrap_add(s.rcmAcc[r], w_diff);
			}
		}
	}
}

/* ConnectivityMatrix.cpp:272-325 */
void applyStdp(oracle_simulation& s, float reward)
{
	if(!s.stdpEnabled) {
		throw oracle_error(NEMO_LOGIC_ERROR, "applyStdp called, but no STDP model specified");
	}
	fix_t fx_reward = fx_toFix(reward, s.fbits);
	if(fx_reward == 0 && reward != 0) {
		throw oracle_error(NEMO_INVALID_INPUT, "STDP reward rounded down to zero");
	}
	if(fx_reward == 0) {
		std::fill(s.rcmAcc.begin(), s.rcmAcc.end(), 0);
	}
	for(size_t r = 0; r < s.rcmAcc.size(); ++r) {
		fix_t& w_old = s.fcm[s.rcmForward[r]].weight;
		fix_t w_new = s.stdp.updatedWeight(w_old, fx_mul(fx_reward, s.rcmAcc[r], s.fbits));
		if(w_old != w_new) {
			w_old = w_new;
		}
		s.rcmAcc[r] = 0;
	}
}

/* cpu/Simulation.cpp:94-111 plus the stimulus half of SimulationBackend::step */
void fire(oracle_simulation& s)
{
	deliverSpikes(s);
	for(auto& g : s.groups) {
		switch(g.type.model) {
			case IZHIKEVICH: updateIzhikevich(s, g); break;
			case QIF: updateQIF(s, g); break;
			case IF_CURR_EXP: updateIFCurrExp(s, g); break;
			case HH: updateHH(s, g); break;
			case POISSON: updatePoisson(s, g); break;
			case INPUT: updateInput(s, g); break;
			case KURAMOTO: case KURAMOTO_LAG: updateKuramoto(s, g); break;
			case IZHIKEVICH_CHAOS: case IZHIKEVICH_SYN: case IZHIKEVICH_2007: case IZHIKEVICH_2007_DOPSYN:
				updateIzhikevichVariant(s, g);
				break;
			case BAYESIAN: updateBayesian(s, g); break;
			case INHOMOGENEOUS_POISSON: updateInhomogeneousPoisson(s, g); break;
		}
	}
	accumulateStdp(s);
	/* setFiring, cpu/Simulation.cpp:200-209: ascending local index */
	s.firedOut.clear();
	for(size_t n = 0; n < s.ncount; ++n) {
		if(s.fired[n]) s.firedOut.push_back(s.l2g[n]);
	}
	s.cycle += 1;
}

/* ------------------------------------------------------------------ */
/* neuron models                                                       */
/* ------------------------------------------------------------------ */

/* cpu/plugins/Izhikevich.cpp:14-98 (parameter indices plugins/Izhikevich.h) */
void updateIzhikevich(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const unsigned start = g.base;
	const size_t stride = g.size;
	const float* a = &g.param[0 * stride];
	const float* b = &g.param[1 * stride];
	const float* c = &g.param[2 * stride];
	const float* d = &g.param[3 * stride];
	const float* sigma = &g.param[4 * stride];
	float* u0 = &g.state[0 * stride];
	float* v0 = &g.state[1 * stride];
	const unsigned SUBSTEPS = 4;
	const float SUBSTEP_MULT = 0.25f;
	const int nn = int(g.size);
	float* currentEPSP = s.currentE.data();
	float* currentIPSP = s.currentI.data();
	float* currentExternal = s.currentExt.data();
	unsigned* fstim = s.fstim.data();
	unsigned* fired = s.fired.data();
	uint64_t* recentFiring = s.recentFiring.data();
	RNG* rng = g.rng.data();

#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		unsigned ng = start + nl;
		float I = currentEPSP[ng] + currentIPSP[ng] + currentExternal[ng];
		currentExternal[ng] = 0.0f;
		if(sigma[nl] != 0.0f) {
			I += sigma[nl] * nrand(&rng[nl]);
		}
		fired[ng] = 0;
		float u = u0[nl];
		float v = v0[nl];
		for(unsigned t = 0; t < SUBSTEPS; ++t) {
			if(!fired[ng]) {
				/* the literals are doubles: this update is evaluated in FP64 */
				v += SUBSTEP_MULT * ((0.04 * v + 5.0) * v + 140.0 - u + I);
				u += SUBSTEP_MULT * (a[nl] * (b[nl] * v - u));
				fired[ng] = v >= 30.0;
			}
		}
		fired[ng] |= fstim[ng];
		fstim[ng] = 0;
		recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
		if(fired[ng]) {
			v = c[nl];
			u += d[nl];
		}
		u0[nl] = u;
		v0[nl] = v;
	}
}

/* cpu/plugins/QIF.cpp:16-107, per-type arrays indexed locally (deviation 1) */
void updateQIF(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const float* a0 = &g.param[0];
	float* v0 = &g.state[0];
	float dt = 0.1f;
	float a1 = 0.5f * dt;
	float a2 = dt;
	float bb1 = dt / 6.0f;
	unsigned int inc_max = (unsigned int)(1 / dt);
	for(unsigned nl = 0; nl < g.size; ++nl) {
		unsigned n = g.base + nl;
		float Excit = s.currentE[n];
		float Inhib = s.currentI[n];
		float Exter = s.currentExt[n];
		s.currentExt[n] = 0.0f;
		s.fired[n] = 0;
		float v = v0[nl];
		float a = a0[nl];
		for(unsigned t = 0; t < inc_max; ++t) {
			if(!s.fired[n]) {
				float I = Excit + Inhib + Exter;
				I = I * 0.015f;
				float k1 = a * v * (v - 1) + I;
				float x2 = v + a1 * k1;
				float k2 = a * x2 * (x2 - 1) + I;
				float x3 = v + a1 * k2;
				float k3 = a * x3 * (x3 - 1) + I;
				float x4 = v + a2 * k3;
				float k4 = a * x4 * (x4 - 1) + I;
				v = v + bb1 * (k1 + 2 * k2 + 2 * k3 + k4);
				s.fired[n] = v >= 1;
			}
		}
		s.fired[n] |= s.fstim[n];
		s.fstim[n] = 0;
		s.recentFiring[n] = (s.recentFiring[n] << 1) | (uint64_t) s.fired[n];
		if(s.fired[n]) {
			v = 0;
		}
		v0[nl] = v;
	}
}

/* cpu/plugins/IF_curr_exp.cpp:27-122 (indices plugins/IF_curr_exp.h:4-18) */
void updateIFCurrExp(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const size_t st = g.size;
	const float* p_v_rest = &g.param[0 * st];
	const float* p_v_reset = &g.param[1 * st];
	const float* p_c_m = &g.param[2 * st];
	const float* p_tau_m = &g.param[3 * st];
	const float* p_tau_refrac = &g.param[4 * st];
	const float* p_tau_syn_E = &g.param[5 * st];
	const float* p_tau_syn_I = &g.param[6 * st];
	const float* p_v_thresh = &g.param[7 * st];
	const float* p_I_offset = &g.param[8 * st];
	float* v0 = &g.state[0 * st];
	float* Ie0 = &g.state[1 * st];
	float* Ii0 = &g.state[2 * st];
	float* lastfired0 = &g.state[3 * st];
	const unsigned start = g.base;
	const int nn = int(g.size);
	float* currentEPSP = s.currentE.data();
	float* currentIPSP = s.currentI.data();
	float* currentExternal = s.currentExt.data();
	unsigned* fstim = s.fstim.data();
	unsigned* fired = s.fired.data();
	uint64_t* recentFiring = s.recentFiring.data();

#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		unsigned ng = start + nl;
		float Ie = ((1.0f - 1.0f / p_tau_syn_E[nl]) * Ie0[nl]) + currentEPSP[ng];
		float Ii = ((1.0f - 1.0f / p_tau_syn_I[nl]) * Ii0[nl]) + currentIPSP[ng];
		float I = Ie + Ii + currentExternal[ng] + p_I_offset[nl];
		Ie0[nl] = Ie;
		Ii0[nl] = Ii;
		currentExternal[ng] = 0.0f;
		float v = v0[nl];
		bool refractory = lastfired0[nl] <= p_tau_refrac[nl];
		if(!refractory) {
			float c_m = p_c_m[nl];
			float v_rest = p_v_rest[nl];
			float tau_m = p_tau_m[nl];
			v += I / c_m + (v_rest - v) / tau_m;
		}
		fired[ng] = (!refractory && v > p_v_thresh[nl]) || fstim[ng];
		fstim[ng] = 0;
		recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
		if(fired[ng]) {
			lastfired0[nl] = 1;
			v0[nl] = p_v_reset[nl];
		} else {
			lastfired0[nl] += 1;
			v0[nl] = v;
		}
	}
}

/* cpu/plugins/HH.cpp:14-138, per-type arrays indexed locally (deviation 1).
 * exp() of a float argument resolves to expf (the <cmath> float overload). */
void updateHH(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const size_t st = g.size;
	float* v0 = &g.state[0 * st];
	float* n0 = &g.state[1 * st];
	float* m0 = &g.state[2 * st];
	float* h0 = &g.state[3 * st];
	float* dir0 = &g.state[4 * st];
	float dt = 0.001f;
	float gNa = 120.0f;
	float gK = 36.0f;
	float gL = 0.3f;
	float ENa = 115.0f - 65.0f;
	float EK = -12.0f - 65.0f;
	float EL = 10.6f - 65.0f;
	float C = 1.0f;
	float RevE = 0.0f;
	float RevI = -70.0f;
	int inc_max = (int)(1 / dt);
	const unsigned long cycle = s.cycle;
	for(unsigned nl = 0; nl < g.size; ++nl) {
		unsigned nn = g.base + nl;
		float v = v0[nl];
		float n = n0[nl];
		float m = m0[nl];
		float h = h0[nl];
		float dir = dir0[nl];
		float Excit = s.currentE[nn];
		float Inhib = s.currentI[nn];
		float Exter = s.currentExt[nn];
		s.currentExt[nn] = 0.0f;
		s.fired[nn] = 0;
		for(int t = 0; t < inc_max; ++t) {
			float I = (Excit * (RevE - v)) + (Inhib * ((RevI - v) / -1)) + Exter;
			float alphan = (0.1f - 0.01f * (v + 65.0f)) / (expf(1.0f - 0.1f * (v + 65.0f)) - 1.0f);
			float alpham = (2.5f - 0.1f * (v + 65.0f)) / (expf(2.5f - 0.1f * (v + 65.0f)) - 1.0f);
			float alphah = 0.07f * expf(-(v + 65.0f) / 20.0f);
			float betan = 0.125f * expf(-(v + 65.0f) / 80.0f);
			float betam = 4.0f * expf(-(v + 65.0f) / 18.0f);
			float betah = 1.0f / (expf(3.0f - 0.1f * (v + 65.0f)) + 1.0f);
			m = m + dt * (alpham * (1.0f - m) - betam * m);
			n = n + dt * (alphan * (1.0f - n) - betan * n);
			h = h + dt * (alphah * (1.0f - h) - betah * h);
			float Ik = gNa * (m * m * m) * h * (v - ENa) + gK * (n * n * n * n) * (v - EK) + gL * (v - EL);
			float newv = v + dt * (-Ik + I) / C;
			float new_dir = (newv - v);
			float change = (dir < 0) | (newv < -45) ? 0 : new_dir;
			dir = new_dir;
			if(!s.fired[nn] && cycle >= 10)
				s.fired[nn] = change < -0.000000001;
			v = newv;
		}
		s.fired[nn] |= s.fstim[nn];
		s.fstim[nn] = 0;
		s.recentFiring[nn] = (s.recentFiring[nn] << 1) | (uint64_t) s.fired[nn];
		v0[nl] = v;
		n0[nl] = n;
		m0[nl] = m;
		h0[nl] = h;
		dir0[nl] = dir;
	}
}

/* cpu/plugins/PoissonSource.cpp:6-30 */
void updatePoisson(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const float* rate = &g.param[0];
	for(unsigned ng = g.base, nl = 0; nl < g.size; ng++, nl++) {
		unsigned p0 = unsigned(rate[nl] * float(1 << 16));
		unsigned p1 = urand(&g.rng[nl]) & 0xffff;
		s.fired[ng] = p1 < p0 || s.fstim[ng];
		s.fstim[ng] = 0;
		s.recentFiring[ng] = (s.recentFiring[ng] << 1) | (uint64_t) s.fired[ng];
	}
}

/* cpu/plugins/Input.cpp:6-26 */
void updateInput(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	for(unsigned n = g.base; n < g.base + g.size; n++) {
		s.fired[n] = s.fstim[n];
		s.fstim[n] = 0;
		s.recentFiring[n] = (s.recentFiring[n] << 1) | (uint64_t) s.fired[n];
	}
}

/* The Izhikevich variants of NeMo's CPU-only plugins (per-type arrays and
 * RNG indexed locally, as every model here):
 *   IzhikevichChaos (cpu/plugins/IzhikevichChaos.cpp:14-101): the Izhikevich
 *     step with 20 substeps of 0.05 ms, and the noise sigma * nrand drawn
 *     every step and added to the u update instead of the input;
 *   IzhikevichSyn (IzhikevichSyn.cpp:14-140, plugins/IzhikevichSyn.h):
 *     conductances ge, gi decaying by a factor per step, incremented by the
 *     excitatory and inhibitory input and clamped to G_exc and -G_inh, input
 *     ge (E_exc - v) - gi (E_inh - v) + external (+ sigma nrand), the
 *     Izhikevich substeps with threshold vpeak;
 *   Izhikevich2007 (Izhikevich2007.cpp:14-134, plugins/Izhikevich2007.h):
 *     C v' = k (v - Vr)(v - Vt) - u + I, u' = a (b (v - Vr) - u) in float,
 *     conductances reset to 1 by any positive input, input
 *     G_exc ge (E_exc - v) + G_inh gi (E_inh - v) + external (+ C sigma nrand);
 *   Izhikevich2007DopSyn (Izhikevich2007DopSyn.cpp:14-156): dopamine levels
 *     d1, d2 scale the excitatory input, Vr, k and d (constants of
 *     plugins/Izhikevich2007DopSyn.h, in double), with IzhikevichSyn's
 *     conductances and input. */
void updateIzhikevichVariant(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const Model m = g.type.model;
	const size_t stride = g.size;
	auto P = [&](unsigned i) { return &g.param[i * stride]; };
	const int nn = int(g.size);
	float* currentEPSP = s.currentE.data();
	float* currentIPSP = s.currentI.data();
	float* currentExternal = s.currentExt.data();
	unsigned* fstim = s.fstim.data();
	unsigned* fired = s.fired.data();
	uint64_t* recentFiring = s.recentFiring.data();
	RNG* rng = g.rng.data();
	float* u0 = &g.state[0 * stride];
	float* v0 = &g.state[1 * stride];
	float* ge0 = g.type.nState > 2 ? &g.state[2 * stride] : nullptr;
	float* gi0 = g.type.nState > 3 ? &g.state[3 * stride] : nullptr;
#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		const unsigned ng = g.base + nl;
		float u = u0[nl];
		float v = v0[nl];
		if(m == IZHIKEVICH_CHAOS) {
			const float *a = P(0), *b = P(1), *c = P(2), *d = P(3), *sigma = P(4);
			const unsigned SUBSTEPS = 20;
			const float SUBSTEP_MULT = 0.05f;
			float I = currentEPSP[ng] + currentIPSP[ng] + currentExternal[ng];
			currentExternal[ng] = 0.0f;
			float randNoise = sigma[nl] * nrand(&rng[nl]);
			fired[ng] = 0;
			for(unsigned t = 0; t < SUBSTEPS; ++t) {
				if(!fired[ng]) {
					v += SUBSTEP_MULT * ((0.04 * v + 5.0) * v + 140.0 - u + I);
					u += SUBSTEP_MULT * (a[nl] * (b[nl] * v - u) + randNoise);
					fired[ng] = v >= 30.0;
				}
			}
			fired[ng] |= fstim[ng];
			fstim[ng] = 0;
			recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
			if(fired[ng]) {
				v = c[nl];
				u += d[nl];
			}
		} else if(m == IZHIKEVICH_SYN) {
			const float *a = P(0), *b = P(1), *c = P(2), *d = P(3), *vpeak = P(4), *sigma = P(5);
			const float *tMdtOt_exc = P(6), *G_exc = P(7), *E_exc = P(8);
			const float *tMdtOt_inh = P(9), *G_inh = P(10), *E_inh = P(11);
			const unsigned SUBSTEPS = 4;
			const float SUBSTEP_MULT = 0.25f;
			float ge = ge0[nl];
			float gi = gi0[nl];
			ge = ge * tMdtOt_exc[nl];
			gi = gi * tMdtOt_inh[nl];
			ge += currentEPSP[ng];
			gi += currentIPSP[ng];
			if(ge > G_exc[nl]) ge = G_exc[nl];
			if(gi < -G_inh[nl]) gi = -G_inh[nl];
			ge0[nl] = ge;
			gi0[nl] = gi;
			float I = ge * (E_exc[nl] - v) - gi * (E_inh[nl] - v) + currentExternal[ng];
			currentExternal[ng] = 0.0f;
			if(sigma[nl] != 0.0f) {
				I += sigma[nl] * nrand(&rng[nl]);
			}
			fired[ng] = 0;
			for(unsigned t = 0; t < SUBSTEPS; ++t) {
				if(!fired[ng]) {
					v += SUBSTEP_MULT * ((0.04 * v + 5.0) * v + 140.0 - u + I);
					u += SUBSTEP_MULT * (a[nl] * (b[nl] * v - u));
					fired[ng] = v >= vpeak[nl];
				}
			}
			fired[ng] |= fstim[ng];
			fstim[ng] = 0;
			recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
			if(fired[ng]) {
				v = c[nl];
				u += d[nl];
			}
		} else {
			const bool dop = m == IZHIKEVICH_2007_DOPSYN;
			const unsigned off = dop ? 2 : 0; /* d1, d2 sit between sigma and the synapse parameters */
			const float *a = P(0), *b = P(1), *c = P(2), *d = P(3), *vpeak = P(4), *Vr = P(5), *Vt = P(6);
			const float *k = P(7), *C = P(8), *sigma = P(9);
			const float *d1 = dop ? P(10) : nullptr, *d2 = dop ? P(11) : nullptr;
			const float *tMdtOt_exc = P(10 + off), *G_exc = P(11 + off), *E_exc = P(12 + off);
			const float *tMdtOt_inh = P(13 + off), *G_inh = P(14 + off), *E_inh = P(15 + off);
			const unsigned SUBSTEPS = 4;
			const float SUBSTEP_MULT = 0.25f;
			const double MSN_BETA1 = 6.3, MSN_BETA2 = 0.215, MSN_K = 0.0289, MSN_L = 0.331, MSN_ALPHA = 0.032;
			float ge = ge0[nl];
			float gi = gi0[nl];
			ge = ge * tMdtOt_exc[nl];
			gi = gi * tMdtOt_inh[nl];
			float I;
			if(dop) {
				ge += (1.0f + MSN_BETA1 * d1[nl]) * (1.0f - MSN_BETA2 * d2[nl]) * currentEPSP[ng];
				gi += currentIPSP[ng];
				if(ge > G_exc[nl]) ge = G_exc[nl];
				if(gi < -G_inh[nl]) gi = -G_inh[nl];
				ge0[nl] = ge;
				gi0[nl] = gi;
				I = ge * (E_exc[nl] - v) - gi * (E_inh[nl] - v) + currentExternal[ng];
			} else {
				if(currentEPSP[ng] > 0.0) ge = 1.0f;
				if(currentIPSP[ng] > 0.0) gi = 1.0f;
				ge0[nl] = ge;
				gi0[nl] = gi;
				I = G_exc[nl] * ge * (E_exc[nl] - v) + G_inh[nl] * gi * (E_inh[nl] - v) + currentExternal[ng];
			}
			currentExternal[ng] = 0.0f;
			if(sigma[nl] != 0.0f) {
				if(dop) I += sigma[nl] * nrand(&rng[nl]);
				else I += C[nl] * sigma[nl] * nrand(&rng[nl]);
			}
			fired[ng] = 0;
			const float oneOverC = 1.0f / C[nl];
			float msn_d = d[nl];
			for(unsigned t = 0; t < SUBSTEPS; ++t) {
				if(!fired[ng]) {
					if(dop) {
						const float msn_Vr = Vr[nl] * (1 + MSN_K * d1[nl]);
						msn_d = d[nl] * (1 - MSN_L * d1[nl]);
						const float msn_k = k[nl] * (1 - MSN_ALPHA * d2[nl]);
						v += SUBSTEP_MULT * oneOverC * (msn_k * (v - msn_Vr) * (v - Vt[nl]) - u + I);
						u += SUBSTEP_MULT * (a[nl] * (b[nl] * (v - msn_Vr) - u));
					} else {
						v += SUBSTEP_MULT * oneOverC * (k[nl] * (v - Vr[nl]) * (v - Vt[nl]) - u + I);
						u += SUBSTEP_MULT * (a[nl] * (b[nl] * (v - Vr[nl]) - u));
					}
					fired[ng] = v >= vpeak[nl];
				}
			}
			fired[ng] |= fstim[ng];
			fstim[ng] = 0;
			recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
			if(fired[ng]) {
				v = c[nl];
				u += dop ? msn_d : d[nl];
			}
		}
		u0[nl] = u;
		v0[nl] = v;
	}
}

/* cpu/plugins/BayesianNeuron.cpp:14-93 (plugins/BayesianNeuron.h: rOn, rOff,
 * g0; state L, G): four Euler substeps of 0.25 of the log-odds L and the
 * adaptation G, with the float exp; fires when L exceeds G + g0/2 (in
 * double), and then G += g0. */
void updateBayesian(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const size_t stride = g.size;
	const float* rOn = &g.param[0 * stride];
	const float* rOff = &g.param[1 * stride];
	const float* g0 = &g.param[2 * stride];
	float* L0 = &g.state[0 * stride];
	float* G0 = &g.state[1 * stride];
	const unsigned SUBSTEPS = 4;
	const float SUBSTEP_MULT = 0.25f;
	const int nn = int(g.size);
#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		const unsigned ng = g.base + nl;
		float I = s.currentE[ng] + s.currentI[ng] + s.currentExt[ng];
		s.currentExt[ng] = 0.0f;
		s.fired[ng] = 0;
		float L = L0[nl];
		float G = G0[nl];
		const float rOn_this = rOn[nl];
		const float rOff_this = rOff[nl];
		const float g0_this = g0[nl];
		for(unsigned t = 0; t < SUBSTEPS; ++t) {
			if(!s.fired[ng]) {
				/* exp of a float is the float overload (libstdc++ <math.h>) */
				L += SUBSTEP_MULT * (rOn_this * (1 + std::exp(-L)) - rOff_this * (1 + std::exp(L)) + I);
				G += SUBSTEP_MULT * (rOn_this * (1 + std::exp(-G)) - rOff_this * (1 + std::exp(G)));
				s.fired[ng] = L > (G + g0_this / 2.0);
			}
		}
		s.fired[ng] |= s.fstim[ng];
		s.fstim[ng] = 0;
		s.recentFiring[ng] = (s.recentFiring[ng] << 1) | (uint64_t) s.fired[ng];
		if(s.fired[ng]) {
			G += g0_this;
		}
		L0[nl] = L;
		G0[nl] = G;
	}
}

/* cpu/plugins/InhomogeneousPoissonSource.cpp:6-27: the firing rate (Hz) is
 * the neuron's external current, which the model reads but does not clear
 * (it holds from step to step until the current stimulus changes it);
 * indexed by the neuron's simulation index (the reference indexes the
 * current by the local one, correct when the type comes first). */
void updateInhomogeneousPoisson(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	for(unsigned ng = g.base, nl = 0; nl < g.size; ng++, nl++) {
		const float rate = s.currentExt[ng];
		unsigned p0 = unsigned(0.001 * rate * float(1 << 16));
		unsigned p1 = urand(&g.rng[nl]) & 0xffff;
		s.fired[ng] = p1 < p0 || s.fstim[ng];
		s.fstim[ng] = 0;
		s.recentFiring[ng] = (s.recentFiring[ng] << 1) | (uint64_t) s.fired[ng];
	}
}

/* cpu/plugins/Kuramoto.cpp:14-134 (parameters plugins/Kuramoto.h: frequency,
 * mean coupling; state: the phase, 32 steps of history), and KuramotoLag
 * (cpu/plugins/KuramotoLag.cpp:15-138, plugins/KuramotoLag.h: a third
 * parameter alpha, subtracted inside every sine: sin(source - target - alpha)).  The phase of each
 * incoming synapse's source is read d-1 steps back; per-type arrays indexed
 * locally (deviation 1), every source being an oscillator of the same
 * group.  The source phase in the reference is taken from slot
 * (cycle - (delay-1)) mod 32 of the unsigned cycle. */
void updateKuramoto(oracle_simulation& s, oracle_simulation::TypeGroup& g)
{
	const unsigned H = g.type.history;
	const unsigned cycle = unsigned(s.cycle);
	const size_t stride = size_t(g.type.nState) * g.size; /* one history slot */
	const float* frequency = &g.param[0];
	const float* cmean = &g.param[g.size];
	const bool lag = g.type.model == KURAMOTO_LAG;
	const float* phase0 = &g.state[size_t(cycle % H) * stride];
	float* phase1 = &g.state[size_t((cycle + 1) % H) * stride];
	g.current = (cycle + 1) % H; /* cpu/Neurons.cpp:123 */
	std::vector<float> weight, sourcePhase;
	for(unsigned nl = 0; nl < g.size; ++nl) {
		const unsigned n = g.base + nl;
		const float h = 0.05f;
		const float f = frequency[nl];
		float targetPhase = phase0[nl];
		/* loadIncoming, Kuramoto.cpp:52-78 */
		const size_t r0 = s.rcmAllStart[n], r1 = s.rcmAllStart[n + 1];
		weight.resize(r1 - r0);
		sourcePhase.resize(r1 - r0);
		for(size_t r = r0; r < r1; ++r) {
			const unsigned src = s.rcmAllSource[r];
			if(src < g.base || src >= g.base + g.size) {
				throw oracle_error(NEMO_API_UNSUPPORTED, "a Kuramoto oscillator's sources must be oscillators of its type");
			}
			weight[r - r0] = s.rcmAllWeight[r];
			const unsigned slot = (cycle - unsigned(int(s.rcmAllDelay[r] - 1))) % H;
			sourcePhase[r - r0] = g.state[size_t(slot) * stride + (src - g.base)];
		}
		/* sumN, Kuramoto.cpp:22-34 */
		const float alpha = lag ? g.param[2 * size_t(g.size) + nl] : 0.0f;
		auto sumN = [&](float target) {
			float sum = 0.0f;
			for(size_t k = 0; k < weight.size(); ++k) {
				sum += weight[k] * (lag ? sinf(sourcePhase[k] - target - alpha) : sinf(sourcePhase[k] - target));
			}
			return sum;
		};
		const float Cmean = cmean[nl] > 0 ? cmean[nl] : 1;
		const float k0 = f + (sumN(targetPhase) / Cmean);
		const float k1 = f + (sumN(targetPhase + k0 * 0.5f * h) / Cmean);
		const float k2 = f + (sumN(targetPhase + k1 * 0.5f * h) / Cmean);
		const float k3 = f + (sumN(targetPhase + k2 * h) / Cmean);
		targetPhase += h * (k0 + 2 * k1 + 2 * k2 + k3) / 6.0f;
		/* float fmodf with 2*pi rounded to float, then a double add */
		phase1[nl] = float(double(fmodf(targetPhase, float(2.0f * M_PI)))
				+ (targetPhase < 0.0f ? 2.0f * M_PI : 0.0f));
	}
}

/* cpu/plugins/Kuramoto.cpp:137-159: the history before the first step, by
 * running each oscillator backwards one frequency (not h times it) per step
 * without coupling, from slot 0 into slots 31, 30, .., 1 */
void initKuramoto(oracle_simulation::TypeGroup& g)
{
	const unsigned H = g.type.history;
	const size_t stride = size_t(g.type.nState) * g.size;
	for(unsigned t = 0; t + 1 < H; ++t) {
		const unsigned current = 0u - t;
		const unsigned previous = 0u - (t + 1u);
		const float* phase0 = &g.state[size_t(current % H) * stride];
		float* phase1 = &g.state[size_t(previous % H) * stride];
		for(unsigned nl = 0; nl < g.size; ++nl) {
			const float phase = phase0[nl] - g.param[nl];
			phase1[nl] = float(double(fmodf(phase, float(2.0f * M_PI))) + (phase < 0.0f ? 2.0f * M_PI : 0.0f));
		}
	}
}

template<typename F>
int guarded(F f)
{
	try {
		f();
		return NEMO_OK;
	} catch(oracle_error& e) {
		g_lastError = e.what();
		return e.code;
	} catch(std::exception& e) {
		g_lastError = e.what();
		return NEMO_UNKNOWN_ERROR;
	}
}

oracle_simulation::TypeGroup& groupOf(oracle_simulation* s, unsigned l)
{
	return s->groups.at(s->typeOfLocal.at(l));
}

} // namespace

/* ------------------------------------------------------------------ */
/* C interface                                                         */
/* ------------------------------------------------------------------ */

extern "C" {

const char* oracle_strerror(void) { return g_lastError.c_str(); }

oracle_network_t oracle_new_network(void) { return new oracle_network(); }
void oracle_delete_network(oracle_network_t n) { delete n; }

int oracle_add_neuron_type(oracle_network_t net, const char* name, unsigned* type)
{
	return guarded([&] {
		auto it = net->typeIds.find(name);
		if(it != net->typeIds.end()) {
			*type = it->second;
			return;
		}
		NeuronType t = lookupType(name);
		unsigned id = unsigned(net->types.size());
		net->types.push_back(t);
		oracle_network::Group g;
		g.param.resize(t.nParam);
		g.state.resize(t.nState);
		net->groups.push_back(g);
		net->typeIds[name] = id;
		*type = id;
	});
}

int oracle_add_neurons(oracle_network_t net, unsigned type, size_t n,
		const unsigned* idx, unsigned nargs, const float* args)
{
	return guarded([&] {
		if(type >= net->types.size()) {
			throw oracle_error(NEMO_INVALID_INPUT, "Invalid neuron type id");
		}
		const NeuronType& t = net->types[type];
		if(nargs != t.nParam + t.nState) {
			throw oracle_error(NEMO_INVALID_INPUT, "Unexpected number of parameters/state variables when adding neuron");
		}
		auto& g = net->groups[type];
		for(size_t i = 0; i < n; ++i) {
			if(!net->mapper.emplace(idx[i], std::make_pair(type, unsigned(g.gidx.size()))).second) {
				throw oracle_error(NEMO_INVALID_INPUT, "Duplicate neuron index " + std::to_string(idx[i]));
			}
			const float* a = args + i * nargs;
			for(unsigned p = 0; p < t.nParam; ++p) g.param[p].push_back(*a++);
			for(unsigned v = 0; v < t.nState; ++v) g.state[v].push_back(*a++);
			g.gidx.push_back(idx[i]);
		}
	});
}

int oracle_add_synapses(oracle_network_t net, size_t n,
		const unsigned* source, const unsigned* target, const unsigned* delay,
		const float* weight, const unsigned char* plastic)
{
	return guarded([&] {
		oracle_network::Axon* last = nullptr;
		unsigned lastSrc = 0;
		for(size_t i = 0; i < n; ++i) {
			if(delay[i] < 1) {
				throw oracle_error(NEMO_INVALID_INPUT, "Invalid delay"); /* NetworkImpl.cpp:98-101 */
			}
			if(!last || source[i] != lastSrc) {
				last = &net->fcm[source[i]];
				lastSrc = source[i];
			}
			auto& ax = *last;
			ax.target.push_back(target[i]);
			ax.delay.push_back(delay[i]);
			ax.weight.push_back(weight[i]);
			ax.plastic.push_back(plastic ? plastic[i] : 0);
			net->maxDelay = std::max(net->maxDelay, delay[i]);
			net->synapseCount += 1;
		}
	});
}

int oracle_neuron_count(oracle_network_t net, unsigned* count)
{
	size_t c = 0;
	for(auto& g : net->groups) c += g.gidx.size();
	*count = unsigned(c);
	return NEMO_OK;
}

oracle_configuration_t oracle_new_configuration(void) { return new oracle_configuration(); }
void oracle_delete_configuration(oracle_configuration_t c) { delete c; }

int oracle_set_stdp_function(oracle_configuration_t conf,
		const float* prefire, size_t prefire_len,
		const float* postfire, size_t postfire_len,
		float minE, float maxE, float minI, float maxI)
{
	return guarded([&] {
		/* StdpFunction.cpp:12-48 */
		if(maxI > 0.0f || minI > 0.0f) {
			throw oracle_error(NEMO_INVALID_INPUT, "STDP function should not have positive limits for inhibitory synapses");
		}
		if(maxE < 0.0f || minE < 0.0f) {
			throw oracle_error(NEMO_INVALID_INPUT, "STDP function should not have negative limits for excitatory synapses");
		}
		if(maxE < minE) {
			throw oracle_error(NEMO_INVALID_INPUT, "STDP should have maximum excitatory weight >= minimum excitatory weight");
		}
		if(maxI > minI) {
			throw oracle_error(NEMO_INVALID_INPUT, "STDP should have abs(maximum inhibitory weight) >= abs(minimum inhibitory weight)");
		}
		if(prefire_len + postfire_len > 64) {
			throw oracle_error(NEMO_INVALID_INPUT, "size of STDP window too large");
		}
		conf->stdp = true;
		conf->prefire.assign(prefire, prefire + prefire_len);
		conf->postfire.assign(postfire, postfire + postfire_len);
		conf->minE = minE;
		conf->maxE = maxE;
		conf->minI = minI;
		conf->maxI = maxI;
	});
}

int oracle_new_simulation(oracle_network_t net, oracle_configuration_t conf, oracle_simulation_t* out)
{
	*out = nullptr;
	return guarded([&] {
		std::unique_ptr<oracle_simulation> s(new oracle_simulation());
		build(*s, *net, *conf);
		*out = s.release();
	});
}

void oracle_delete_simulation(oracle_simulation_t s) { delete s; }

int oracle_step(oracle_simulation_t s,
		const unsigned* fstim, size_t fstim_count,
		const unsigned* istim_nidx, const float* istim_current, size_t istim_count,
		const unsigned** fired, size_t* fired_count)
{
	return guarded([&] {
		/* SimulationBackend.cpp:61-77: current stimulus, then firing stimulus
		 * (cpu/Simulation.cpp:128-156) */
		for(size_t i = 0; i < istim_count; ++i) {
			s->currentExt.at(s->localIdx(istim_nidx[i])) = istim_current[i];
		}
		for(size_t i = 0; i < fstim_count; ++i) {
			s->fstim.at(s->localIdx(fstim[i])) = 1;
		}
		fire(*s);
		if(fired) *fired = s->firedOut.data();
		if(fired_count) *fired_count = s->firedOut.size();
	});
}

int oracle_apply_stdp(oracle_simulation_t s, float reward)
{
	return guarded([&] { applyStdp(*s, reward); });
}

int oracle_get_membrane_potential(oracle_simulation_t s, unsigned neuron, float* v)
{
	return guarded([&] {
		unsigned l = s->localIdx(neuron);
		auto& g = groupOf(s, l);
		if(g.type.membranePotential >= g.type.nState) {
			throw oracle_error(NEMO_INVALID_INPUT, "neuron type has no membrane potential");
		}
		*v = g.at(g.type.membranePotential, l - g.base);
	});
}

int oracle_get_neuron_state(oracle_simulation_t s, unsigned neuron, unsigned var, float* val)
{
	return guarded([&] {
		unsigned l = s->localIdx(neuron);
		auto& g = groupOf(s, l);
		if(var >= g.type.nState) throw oracle_error(NEMO_INVALID_INPUT, "Invalid state variable index");
		*val = g.at(var, l - g.base);
	});
}

int oracle_get_neuron_parameter(oracle_simulation_t s, unsigned neuron, unsigned param, float* val)
{
	return guarded([&] {
		unsigned l = s->localIdx(neuron);
		auto& g = groupOf(s, l);
		if(param >= g.type.nParam) throw oracle_error(NEMO_INVALID_INPUT, "Invalid parameter index");
		*val = g.param[size_t(param) * g.size + (l - g.base)];
	});
}

int oracle_set_neuron_state(oracle_simulation_t s, unsigned neuron, unsigned var, float val)
{
	return guarded([&] {
		unsigned l = s->localIdx(neuron);
		auto& g = groupOf(s, l);
		if(var >= g.type.nState) throw oracle_error(NEMO_INVALID_INPUT, "Invalid state variable index");
		g.at(var, l - g.base) = val;
	});
}

int oracle_set_neuron_parameter(oracle_simulation_t s, unsigned neuron, unsigned param, float val)
{
	return guarded([&] {
		unsigned l = s->localIdx(neuron);
		auto& g = groupOf(s, l);
		if(param >= g.type.nParam) throw oracle_error(NEMO_INVALID_INPUT, "Invalid parameter index");
		g.param[size_t(param) * g.size + (l - g.base)] = val;
	});
}

static std::vector<uint64_t> g_queried;

int oracle_get_synapses_from(oracle_simulation_t s, unsigned source, const uint64_t** ids, size_t* len)
{
	return guarded([&] {
		/* ConnectivityMatrix.cpp:329-364: ids are (source << 32) | index */
		g_queried.clear();
		auto it = s->auxRange.find(source);
		if(it == s->auxRange.end()) {
			s->localIdx(source);
		} else {
			for(size_t k = 0; k < it->second.second - it->second.first; ++k) {
				g_queried.push_back((uint64_t(source) << 32) | uint64_t(k));
			}
		}
		*ids = g_queried.data();
		*len = g_queried.size();
	});
}

static size_t synapseSeq(oracle_simulation_t s, uint64_t id)
{
	unsigned src = unsigned(id >> 32);
	size_t k = size_t(id & 0xffffffffu);
	auto it = s->auxRange.find(src);
	if(it == s->auxRange.end() || k >= it->second.second - it->second.first) {
		throw oracle_error(NEMO_INVALID_INPUT, "Non-existing synapse id");
	}
	return it->second.first + k;
}

int oracle_get_synapse_target(oracle_simulation_t s, uint64_t id, unsigned* target)
{
	return guarded([&] { *target = s->l2g.at(s->fcm[s->auxAddr[synapseSeq(s, id)]].target); });
}

int oracle_get_synapse_delay(oracle_simulation_t s, uint64_t id, unsigned* delay)
{
	return guarded([&] { *delay = s->auxDelay[synapseSeq(s, id)]; });
}

int oracle_get_synapse_weight(oracle_simulation_t s, uint64_t id, float* weight)
{
	return guarded([&] { *weight = fx_toFloat(s->fcm[s->auxAddr[synapseSeq(s, id)]].weight, s->fbits); });
}

int oracle_get_synapse_plastic(oracle_simulation_t s, uint64_t id, unsigned char* plastic)
{
	return guarded([&] { *plastic = s->auxPlastic[synapseSeq(s, id)]; });
}

int oracle_elapsed_simulation(oracle_simulation_t s, unsigned long* ms)
{
	*ms = s->cycle;
	return NEMO_OK;
}

int oracle_reset_timer(oracle_simulation_t s)
{
	s->cycle = 0;
	return NEMO_OK;
}

int oracle_local_to_global(oracle_simulation_t s, unsigned* gidx, size_t n)
{
	return guarded([&] {
		if(n != s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		std::copy(s->l2g.begin(), s->l2g.end(), gidx);
	});
}

int oracle_get_currents(oracle_simulation_t s, float* E, float* I, size_t n)
{
	return guarded([&] {
		if(n != s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		std::copy(s->currentE.begin(), s->currentE.end(), E);
		std::copy(s->currentI.begin(), s->currentI.end(), I);
	});
}

int oracle_get_fixed_weights(oracle_simulation_t s, int32_t* w, size_t n)
{
	return guarded([&] {
		if(n != s->auxAddr.size()) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		for(size_t k = 0; k < n; ++k) w[k] = s->fcm[s->auxAddr[k]].weight;
	});
}

int oracle_get_stdp_accumulators(oracle_simulation_t s, int32_t* acc, size_t n)
{
	return guarded([&] {
		if(n != s->auxAddr.size()) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		std::vector<int32_t> byAddr(s->fcm.size(), 0);
		for(size_t r = 0; r < s->rcmAcc.size(); ++r) byAddr[s->rcmForward[r]] = s->rcmAcc[r];
		for(size_t k = 0; k < n; ++k) acc[k] = byAddr[s->auxAddr[k]];
	});
}

int oracle_get_recent_firing(oracle_simulation_t s, uint64_t* hist, size_t n)
{
	return guarded([&] {
		if(n != s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		std::copy(s->recentFiring.begin(), s->recentFiring.end(), hist);
	});
}

int oracle_get_rng_state(oracle_simulation_t s, uint32_t* state, size_t n)
{
	return guarded([&] {
		if(n != s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		for(auto& g : s->groups) {
			for(unsigned i = 0; i < g.size; ++i) {
				std::memcpy(state + 4 * size_t(g.base + i), g.rng[i].state, 16);
			}
		}
	});
}

uint64_t oracle_delivered_events(oracle_simulation_t s) { return s->delivered; }

int oracle_num_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

void oracle_rng_seeds(unsigned first, unsigned last, uint32_t* out)
{
	std::vector<RNG> r(last - first + 1);
	initialiseRng(first, last, r);
	std::memcpy(out, r.data(), r.size() * sizeof(RNG));
}

uint32_t oracle_urand(uint32_t* state4)
{
	RNG r;
	std::memcpy(r.state, state4, 16);
	uint32_t v = urand(&r);
	std::memcpy(state4, r.state, 16);
	return v;
}

float oracle_nrand(uint32_t* state4)
{
	RNG r;
	std::memcpy(r.state, state4, 16);
	float v = nrand(&r);
	std::memcpy(state4, r.state, 16);
	return v;
}

/* state variable var of every neuron in local order (NaN where its model has
 * fewer) */
int oracle_get_neuron_states(oracle_simulation_t s, unsigned var, float* out, size_t n)
{
	return guarded([&] {
		if(n != s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "size mismatch");
		for(size_t l = 0; l < n; ++l) {
			oracle_simulation::TypeGroup& g = groupOf(s, unsigned(l));
			out[l] = var < g.type.nState ? g.at(var, unsigned(l - g.base)) : std::nanf("");
		}
	});
}

/* cpu/Simulation.cpp:115-137 (propagate_raw, NEMO_BRIAN_ENABLED): every
 * history shifts by one, the given neurons (local indices) fire, and the
 * spikes arriving now are delivered; the timer advances. */
int oracle_propagate(oracle_simulation_t s, const unsigned* fired, size_t nfired, const float** E, const float** I)
{
	return guarded([&] {
		for(size_t n = 0; n < s->ncount; ++n) s->recentFiring[n] <<= 1;
		for(size_t i = 0; i < nfired; ++i) {
			if(fired[i] >= s->ncount) throw oracle_error(NEMO_INVALID_INPUT, "propagate: neuron index out of range");
			s->recentFiring[fired[i]] |= uint64_t(1);
		}
		deliverSpikes(*s);
		s->cycle += 1;
		*E = s->currentE.data();
		*I = s->currentI.data();
	});
}

void oracle_nrand_n(uint32_t* states, size_t n, float* out)
{
	for(size_t i = 0; i < n; ++i) out[i] = oracle_nrand(states + 4 * i);
}

uint32_t oracle_mt19937_nth(uint32_t seed, unsigned n)
{
	std::mt19937 mt(seed);
	uint32_t v = 0;
	for(unsigned i = 0; i < n; ++i) v = uint32_t(mt());
	return v;
}

int oracle_fx_tofix(float f, unsigned fbits, int32_t* out)
{
	return guarded([&] { *out = fx_toFix(f, fbits); });
}

float oracle_wfx_tofloat(int64_t v, unsigned fbits) { return wfx_toFloat(v, fbits); }

} // extern "C"
