/*
 * nemo_c.cpp -- the nemo.h C API over the HIP backend. Exceptions become
 * status codes plus a process-wide message, as NeMo's src/nemo/nemo_c.cpp:
 * 24-65 does.
 */
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "nemo.h"
#include "nemo_hip.h"
#include "common.hpp"
#include "network.hpp"
#include "simulation.hpp"
#include "batch.hpp"

using nemo::hip::NetworkImpl;
using nemo::hip::ConfigurationImpl;
using nemo::hip::Simulation;

struct nemo_network_class { NetworkImpl impl; };
struct nemo_configuration_class { ConfigurationImpl impl; };
struct nemo_simulation_class {
	Simulation* sim;
	std::vector<unsigned> fired;
};

static std::string g_lastError;
static nemo_status_t g_lastStatus = NEMO_OK;

#define CALL(...) do { \
	g_lastStatus = NEMO_OK; \
	try { __VA_ARGS__; } \
	catch(nemo::exception& e) { g_lastError = e.what(); g_lastStatus = e.errorNumber(); } \
	catch(std::bad_alloc& e) { g_lastError = "host allocation failed"; g_lastStatus = NEMO_ALLOCATION_ERROR; } \
	catch(std::exception& e) { g_lastError = e.what(); g_lastStatus = NEMO_UNKNOWN_ERROR; } \
	catch(...) { g_lastError = "unknown exception"; g_lastStatus = NEMO_UNKNOWN_ERROR; } \
	} while(0)

#define RETURN_CALL(...) do { CALL(__VA_ARGS__); return g_lastStatus; } while(0)

static void
requireSim(nemo_simulation_t s)
{
	if(!s || !s->sim) throw nemo::exception(NEMO_INVALID_INPUT, "null simulation");
}

extern "C" {

const char* nemo_version(void) { return "0.7.2-hip"; }

nemo_status_t
nemo_add_plugin_path(const char* dir)
{
	RETURN_CALL(
		if(!dir) throw nemo::exception(NEMO_INVALID_INPUT, "null plugin path");
		nemo::addPluginPath(std::string(dir));
	);
}

nemo_status_t
nemo_cuda_device_count(unsigned* count)
{
	RETURN_CALL(
		int n = 0;
		hipError_t e = hipGetDeviceCount(&n);
		if(e != hipSuccess) n = 0;
		*count = unsigned(n);
		if(e != hipSuccess) nemo::hipCheck(e, "hipGetDeviceCount");
	);
}

nemo_status_t
nemo_cuda_device_description(unsigned device, const char** descr)
{
	static std::string s;
	RETURN_CALL(
		hipDeviceProp_t p;
		nemo::hipCheck(hipGetDeviceProperties(&p, int(device)), "hipGetDeviceProperties");
		s = nemo::fmt(p.name, " (", p.gcnArchName, ", ", p.multiProcessorCount, " CUs, ",
				(p.totalGlobalMem >> 20), " MiB)");
		*descr = s.c_str();
	);
}

/* --------------------------------------------------------------- configuration */

nemo_configuration_t nemo_new_configuration(void) { return new nemo_configuration_class(); }
void nemo_delete_configuration(nemo_configuration_t c) { delete c; }

nemo_status_t nemo_log_stdout(nemo_configuration_t c) { RETURN_CALL(c->impl.logging = true); }

nemo_status_t
nemo_set_stdp_function(nemo_configuration_t c, float pre[], size_t npre, float post[], size_t npost,
		float minE, float maxE, float minI, float maxI)
{
	RETURN_CALL(
		c->impl.setStdpFunction(std::vector<float>(pre, pre + npre), std::vector<float>(post, post + npost),
				minE, maxE, minI, maxI);
	);
}

nemo_status_t
nemo_set_cpu_backend(nemo_configuration_t c)
{
	RETURN_CALL(
		(void) c;
		throw nemo::exception(NEMO_API_UNSUPPORTED,
				"libnemo_hip provides the GPU backend only; NeMo's CPU backend is libnemo_cpu");
	);
}

nemo_status_t
nemo_set_cuda_backend(nemo_configuration_t c, int dev)
{
	RETURN_CALL(
		int n = 0;
		nemo::hipCheck(hipGetDeviceCount(&n), "hipGetDeviceCount");
		int d = dev < 0 ? 0 : dev;
		if(d >= n) throw nemo::exception(NEMO_INVALID_INPUT, nemo::fmt("Invalid HIP device ", dev));
		c->impl.backend = NEMO_BACKEND_CUDA;
		c->impl.device = d;
	);
}

nemo_status_t nemo_set_hip_backend(nemo_configuration_t c, int dev) { return nemo_set_cuda_backend(c, dev); }

/* ---- accessors behind the C++ surface (nemo.hpp) that nemo.h lacks */

nemo_configuration_t
nemo_copy_configuration(nemo_configuration_t c)
{
	nemo_configuration_t out = nullptr;
	CALL(
		if(!c) throw nemo::exception(NEMO_INVALID_INPUT, "null configuration");
		out = new nemo_configuration_class(*c);
	);
	return out;
}

nemo_status_t nemo_disable_logging(nemo_configuration_t c) { RETURN_CALL(c->impl.logging = false); }
nemo_status_t nemo_logging_enabled(nemo_configuration_t c, int* on) { RETURN_CALL(*on = c->impl.logging ? 1 : 0); }
nemo_status_t nemo_partition_size(nemo_configuration_t c, unsigned* ps) { RETURN_CALL(*ps = c->impl.partitionSize); }
nemo_status_t nemo_write_only_synapses(nemo_configuration_t c, int* on) { RETURN_CALL(*on = c->impl.writeOnlySynapses ? 1 : 0); }

nemo_status_t nemo_cuda_device(nemo_configuration_t c, int* dev) { RETURN_CALL(*dev = c->impl.device); }

nemo_status_t nemo_backend(nemo_configuration_t c, backend_t* b) { RETURN_CALL(*b = backend_t(c->impl.backend)); }

nemo_status_t
nemo_backend_description(nemo_configuration_t c, const char** descr)
{
	RETURN_CALL(
		c->impl.description = nemo::fmt("HIP backend (device ", c->impl.device, ")");
		*descr = c->impl.description.c_str();
	);
}

nemo_status_t nemo_set_write_only_synapses(nemo_configuration_t c) { RETURN_CALL(c->impl.writeOnlySynapses = true); }

nemo_status_t nemo_set_partition_size(nemo_configuration_t c, unsigned n) { RETURN_CALL(c->impl.partitionSize = n); }

nemo_status_t
nemo_set_delivery_lanes(nemo_configuration_t c, unsigned lanes)
{
	RETURN_CALL(
		if(lanes > 64 || (lanes & (lanes - 1)) != 0) {
			throw nemo::exception(NEMO_INVALID_INPUT, nemo::fmt("delivery lanes must be 0 or a power of two up to 64, not ", lanes));
		}
		c->impl.deliveryLanes = lanes;
	);
}

nemo_status_t nemo_set_stdp_lds_staging(nemo_configuration_t c, int enable) { RETURN_CALL(c->impl.stdpLdsStaging = enable != 0); }
nemo_status_t nemo_set_exchange_waves(nemo_configuration_t c, int waves) { RETURN_CALL(c->impl.exchangeWaves = waves); }

nemo_status_t nemo_set_device_construction(nemo_configuration_t c, int enable) { RETURN_CALL(c->impl.deviceConstruction = enable != 0); }

nemo_status_t
nemo_set_step_mode(nemo_configuration_t c, int mode)
{
	RETURN_CALL(
		if(mode != NEMO_STEP_AUTO && mode != NEMO_STEP_FUSED && mode != NEMO_STEP_PLUGINS) {
			throw nemo::exception(NEMO_INVALID_INPUT, nemo::fmt("Invalid step mode ", mode));
		}
		c->impl.stepMode = mode;
	);
}

/* ----------------------------------------------------------------- construction */

nemo_network_t nemo_new_network(void) { return new nemo_network_class(); }
void nemo_delete_network(nemo_network_t n) { delete n; }

nemo_status_t
nemo_add_neuron_type(nemo_network_t net, const char* name, unsigned* type)
{
	RETURN_CALL(*type = net->impl.addNeuronType(std::string(name)));
}

/* src/nemo/Network.cpp:46-56: argument order {a, b, c, d, sigma, u, v} */
nemo_status_t
nemo_add_neuron_iz(nemo_network_t net, unsigned idx, float a, float b, float c, float d,
		float u, float v, float sigma)
{
	RETURN_CALL(
		unsigned t = net->impl.addNeuronType("Izhikevich");
		float args[7] = { a, b, c, d, sigma, u, v };
		net->impl.addNeuron(t, idx, 7, args);
	);
}

nemo_status_t
nemo_add_neuron(nemo_network_t net, unsigned type, unsigned idx, unsigned nargs, float args[])
{
	RETURN_CALL(net->impl.addNeuron(type, idx, nargs, args));
}

nemo_status_t
nemo_add_neurons(nemo_network_t net, unsigned type, size_t n, const unsigned idx[], unsigned nargs, const float args[])
{
	RETURN_CALL(net->impl.addNeurons(type, n, idx, nargs, args));
}

nemo_status_t
nemo_add_synapse(nemo_network_t net, unsigned source, unsigned target, unsigned delay, float weight,
		unsigned char plastic, synapse_id* id)
{
	RETURN_CALL(
		synapse_id i = net->impl.addSynapse(source, target, delay, weight, plastic);
		if(id) *id = i;
	);
}

nemo_status_t
nemo_add_synapses_csr(nemo_network_t net, size_t nsources, const unsigned sources[], const uint64_t rowptr[],
		const unsigned targets[], const unsigned char delays[], const float weights[], const unsigned char plastic[])
{
	RETURN_CALL(net->impl.addSynapsesCsr(nsources, sources, rowptr, targets, delays, weights, plastic));
}

nemo_status_t nemo_neuron_count(nemo_network_t net, unsigned* n) { RETURN_CALL(*n = net->impl.neuronCount()); }

nemo_status_t nemo_max_delay(nemo_network_t net, unsigned* d) { RETURN_CALL(*d = net->impl.maxDelay()); }

/* src/nemo/NetworkImpl.cpp:28-30,125-127: both start at 0 and follow every
 * added synapse */
nemo_status_t
nemo_min_weight(nemo_network_t net, float* w)
{
	RETURN_CALL(
		float m = 0.0f;
		for(float x : net->impl.weight()) m = std::min(m, x);
		*w = m;
	);
}

nemo_status_t
nemo_max_weight(nemo_network_t net, float* w)
{
	RETURN_CALL(
		float m = 0.0f;
		for(float x : net->impl.weight()) m = std::max(m, x);
		*w = m;
	);
}

/* ------------------------------------------------------------------- simulation */

nemo_simulation_t
nemo_new_simulation(nemo_network_t net, nemo_configuration_t conf)
{
	nemo_simulation_t s = nullptr;
	CALL(
		if(!net || !conf) throw nemo::exception(NEMO_INVALID_INPUT, "null network or configuration");
		if(conf->impl.backend != NEMO_BACKEND_CUDA) throw nemo::exception(NEMO_API_UNSUPPORTED, "only the HIP backend is available");
		Simulation* sim = new Simulation(net->impl, conf->impl);
		s = new nemo_simulation_class();
		s->sim = sim;
	);
	return s;
}

void
nemo_delete_simulation(nemo_simulation_t s)
{
	if(s) {
		delete s->sim;
		delete s;
	}
}

nemo_status_t
nemo_step(nemo_simulation_t s, unsigned fstim[], size_t nf, unsigned istim_idx[], float istim_val[], size_t ni,
		unsigned* fired[], size_t* nfired)
{
	RETURN_CALL(
		requireSim(s);
		std::vector<unsigned> fs(fstim, fstim + nf);
		const std::vector<unsigned>& f = s->sim->step(fs, istim_idx, istim_val, ni);
		s->fired = f;
		if(fired) *fired = s->fired.empty() ? nullptr : s->fired.data();
		if(nfired) *nfired = s->fired.size();
	);
}

nemo_status_t nemo_apply_stdp(nemo_simulation_t s, float reward) { RETURN_CALL(requireSim(s); s->sim->applyStdp(reward)); }

nemo_status_t
nemo_hip_apply_stdp_batch(nemo_simulation_t sims[], size_t n, const float rewards[])
{
	RETURN_CALL(
		if(n == 0) return NEMO_OK;
		if(!sims || !rewards) throw nemo::exception(NEMO_INVALID_INPUT, "null simulation or reward array");
		for(size_t i = 0; i < n; ++i) requireSim(sims[i]);
		std::lock_guard<nemo::DeviceMutex> devLock(nemo::deviceMutex());
		/* all launches first, then all waits: the kernels of different
		 * simulations run on their own streams at once */
		size_t launched = 0;
		try {
			for(; launched < n; ++launched) sims[launched]->sim->applyStdpLaunch(rewards[launched]);
		} catch(...) {
			for(size_t i = 0; i < launched; ++i) sims[i]->sim->applyStdpFinish();
			throw;
		}
		for(size_t i = 0; i < n; ++i) sims[i]->sim->applyStdpFinish();
	);
}

nemo_status_t
nemo_hip_run(nemo_simulation_t s, unsigned nsteps, const nemo_schedule_t* sched, unsigned long long* events)
{
	RETURN_CALL(requireSim(s); s->sim->run(nsteps, sched, events));
}

nemo_status_t
nemo_hip_run_static(nemo_simulation_t s, unsigned nsteps, const nemo_schedule_t* sched, uint64_t* raster_dev,
		unsigned long long* events)
{
	RETURN_CALL(requireSim(s); s->sim->runStatic(nsteps, sched, raster_dev, events));
}

nemo_status_t
nemo_hip_read_fired(nemo_simulation_t s, unsigned counts[], unsigned ids[], size_t cap, size_t* total)
{
	RETURN_CALL(requireSim(s); s->sim->readFired(counts, ids, cap, total));
}

nemo_status_t
nemo_hip_read_raster(nemo_simulation_t s, uint64_t words[], size_t cap, size_t* wps)
{
	RETURN_CALL(requireSim(s); s->sim->readRaster(words, cap, wps));
}

nemo_status_t nemo_hip_get_info(nemo_simulation_t s, nemo_hip_info_t* info) { RETURN_CALL(requireSim(s); s->sim->info(info)); }

/* Batched persistent launches: simulations are packed into launches in the
 * given order while their workgroups fit on the device at once; each
 * network keeps its own barrier, so the networks of a launch never wait
 * for each other. */
nemo_status_t
nemo_hip_run_batch(nemo_simulation_t sims[], size_t n, const unsigned nsteps[], unsigned long long events[])
{
	RETURN_CALL(
		if(n == 0) return NEMO_OK;
		if(!sims) throw nemo::exception(NEMO_INVALID_INPUT, "null simulation array");
		/* a simulation listed twice would share its state and exchange
		 * buffers between two sets of workgroups of one launch */
		std::vector<nemo_simulation_t> sorted(sims, sims + n);
		std::sort(sorted.begin(), sorted.end());
		if(std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
			throw nemo::exception(NEMO_INVALID_INPUT, "a simulation appears more than once in the batch");
		}
		std::vector<Simulation*> ss(n);
		for(size_t i = 0; i < n; ++i) {
			requireSim(sims[i]);
			ss[i] = sims[i]->sim;
		}
		nemo::hip::runBatch(ss.data(), n, nsteps, events);
	);
}

/* ---------------------------------------------------------------------- queries */

nemo_status_t nemo_get_membrane_potential(nemo_simulation_t s, unsigned n, float* v) { RETURN_CALL(requireSim(s); *v = s->sim->getMembranePotential(n)); }
nemo_status_t nemo_get_neuron_state_n(nemo_network_t net, unsigned n, unsigned var, float* v) { RETURN_CALL(*v = net->impl.getNeuronState(n, var)); }
nemo_status_t nemo_get_neuron_parameter_n(nemo_network_t net, unsigned n, unsigned p, float* v) { RETURN_CALL(*v = net->impl.getNeuronParameter(n, p)); }
nemo_status_t nemo_get_neuron_state_s(nemo_simulation_t s, unsigned n, unsigned var, float* v) { RETURN_CALL(requireSim(s); *v = s->sim->getNeuronState(n, var)); }
nemo_status_t nemo_get_neuron_parameter_s(nemo_simulation_t s, unsigned n, unsigned p, float* v) { RETURN_CALL(requireSim(s); *v = s->sim->getNeuronParameter(n, p)); }

nemo_status_t nemo_get_synapse_source_n(nemo_network_t net, synapse_id id, unsigned* v) { RETURN_CALL(*v = net->impl.getSynapseSource(id)); }
nemo_status_t nemo_get_synapse_source_s(nemo_simulation_t s, synapse_id id, unsigned* v) { RETURN_CALL(requireSim(s); *v = s->sim->getSynapseSource(id)); }
nemo_status_t nemo_get_synapse_target_n(nemo_network_t net, synapse_id id, unsigned* v) { RETURN_CALL(*v = net->impl.getSynapseTarget(id)); }
nemo_status_t nemo_get_synapse_target_s(nemo_simulation_t s, synapse_id id, unsigned* v) { RETURN_CALL(requireSim(s); *v = s->sim->getSynapseTarget(id)); }
nemo_status_t nemo_get_synapse_delay_n(nemo_network_t net, synapse_id id, unsigned* v) { RETURN_CALL(*v = net->impl.getSynapseDelay(id)); }
nemo_status_t nemo_get_synapse_delay_s(nemo_simulation_t s, synapse_id id, unsigned* v) { RETURN_CALL(requireSim(s); *v = s->sim->getSynapseDelay(id)); }
nemo_status_t nemo_get_synapse_weight_n(nemo_network_t net, synapse_id id, float* v) { RETURN_CALL(*v = net->impl.getSynapseWeight(id)); }
nemo_status_t nemo_get_synapse_weight_s(nemo_simulation_t s, synapse_id id, float* v) { RETURN_CALL(requireSim(s); *v = s->sim->getSynapseWeight(id)); }
nemo_status_t nemo_get_synapse_plastic_n(nemo_network_t net, synapse_id id, unsigned char* v) { RETURN_CALL(*v = net->impl.getSynapsePlastic(id)); }
nemo_status_t nemo_get_synapse_plastic_s(nemo_simulation_t s, synapse_id id, unsigned char* v) { RETURN_CALL(requireSim(s); *v = s->sim->getSynapsePlastic(id)); }

nemo_status_t
nemo_get_synapses_from_n(nemo_network_t net, unsigned source, synapse_id* synapses[], size_t* len)
{
	RETURN_CALL(
		const std::vector<synapse_id>& ids = net->impl.getSynapsesFrom(source);
		*synapses = ids.empty() ? nullptr : const_cast<synapse_id*>(ids.data());
		*len = ids.size();
	);
}

nemo_status_t
nemo_get_synapses_from_s(nemo_simulation_t s, unsigned source, synapse_id* synapses[], size_t* len)
{
	RETURN_CALL(
		requireSim(s);
		const std::vector<synapse_id>& ids = s->sim->getSynapsesFrom(source);
		*synapses = ids.empty() ? nullptr : const_cast<synapse_id*>(ids.data());
		*len = ids.size();
	);
}

/* ------------------------------------------------------------------ modification */

nemo_status_t nemo_set_neuron_n(nemo_network_t net, unsigned idx, unsigned nargs, float args[]) { RETURN_CALL(net->impl.setNeuron(idx, nargs, args)); }
nemo_status_t nemo_set_neuron_s(nemo_simulation_t s, unsigned idx, unsigned nargs, float args[]) { RETURN_CALL(requireSim(s); s->sim->setNeuron(idx, nargs, args)); }

nemo_status_t
nemo_set_neuron_iz_n(nemo_network_t net, unsigned idx, float a, float b, float c, float d, float u, float v, float sigma)
{
	float args[7] = { a, b, c, d, sigma, u, v };
	return nemo_set_neuron_n(net, idx, 7, args);
}

nemo_status_t
nemo_set_neuron_iz_s(nemo_simulation_t s, unsigned idx, float a, float b, float c, float d, float u, float v, float sigma)
{
	float args[7] = { a, b, c, d, sigma, u, v };
	return nemo_set_neuron_s(s, idx, 7, args);
}

nemo_status_t nemo_set_neuron_state_n(nemo_network_t net, unsigned n, unsigned var, float v) { RETURN_CALL(net->impl.setNeuronState(n, var, v)); }
nemo_status_t nemo_set_neuron_parameter_n(nemo_network_t net, unsigned n, unsigned p, float v) { RETURN_CALL(net->impl.setNeuronParameter(n, p, v)); }
nemo_status_t nemo_set_neuron_state_s(nemo_simulation_t s, unsigned n, unsigned var, float v) { RETURN_CALL(requireSim(s); s->sim->setNeuronState(n, var, v)); }
nemo_status_t nemo_set_neuron_parameter_s(nemo_simulation_t s, unsigned n, unsigned p, float v) { RETURN_CALL(requireSim(s); s->sim->setNeuronParameter(n, p, v)); }

/* ------------------------------------------------------------------------ timers */

nemo_status_t nemo_elapsed_wallclock(nemo_simulation_t s, unsigned long* v) { RETURN_CALL(requireSim(s); *v = s->sim->elapsedWallclock()); }
nemo_status_t nemo_elapsed_simulation(nemo_simulation_t s, unsigned long* v) { RETURN_CALL(requireSim(s); *v = s->sim->elapsedSimulation()); }
nemo_status_t nemo_reset_timer(nemo_simulation_t s) { RETURN_CALL(requireSim(s); s->sim->resetTimer()); }

const char* nemo_strerror(void) { return g_lastError.c_str(); }

nemo_status_t nemo_hip_last_status(void) { return g_lastStatus; }

/* ------------------------------------------------------------------------- math */

nemo_status_t
nemo_hip_eval_libm(int kind, const float in[], float out[], size_t n, int on_device)
{
	RETURN_CALL(
		if(kind < 0 || kind > 3) {
			throw nemo::exception(NEMO_INVALID_INPUT,
					"kind must be 0 (logf), 1 (the RNG's sinf), 2 (Kuramoto phase wrap) or 3 (sinf, any argument)");
		}
		if(!on_device) {
			for(size_t i = 0; i < n; ++i) {
				out[i] = kind == 0 ? nemo_libm::logf_glibc(in[i])
					: kind == 1 ? nemo_libm::sinf_rng(in[i])
					: kind == 2 ? nemo_libm::kuramoto_wrap(in[i]) : nemo_libm::sinf_glibc(in[i]);
			}
		} else {
			nemo::hip::DevBuf din, dout;
			din.upload(in, n * 4);
			dout.alloc(n * 4);
			nemo::hipCheck(nemo_dev_eval_libm(kind, din.as<float>(), dout.as<float>(), n, nullptr), "libm kernel");
			nemo::hipCheck(hipMemcpy(out, dout.ptr, n * 4, hipMemcpyDeviceToHost), "D2H");
		}
	);
}

nemo_status_t
nemo_hip_nrand_seq(const uint32_t state4[4], float out[], size_t n)
{
	RETURN_CALL(
		nemo::hip::DevBuf ds, dout;
		ds.upload(state4, 16);
		dout.alloc(std::max<size_t>(n, 1) * 4);
		nemo::hipCheck(nemo_dev_nrand_seq(ds.as<uint32_t>(), dout.as<float>(), n, nullptr), "nrand kernel");
		nemo::hipCheck(hipMemcpy(out, dout.ptr, n * 4, hipMemcpyDeviceToHost), "D2H");
	);
}

} // extern "C"
