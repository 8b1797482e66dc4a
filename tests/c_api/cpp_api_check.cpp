/*
 * cpp_api_check.cpp -- NeMo's C++ API (src/nemo.hpp: nemo::Network,
 * nemo::Configuration, nemo::Simulation) through include/nemo.hpp, built
 * with g++ and linked with libnemo_hip.so, the way a NeMo C++ user would use
 * the library. tests/test_c_api.py runs it.
 *
 *   cpp_api_check net
 *       network and configuration only, no device: neuron count, maximum
 *       delay, weight range (from 0, src/nemo/NetworkImpl.cpp:28-30,125-127),
 *       synapse ids and queries, parameters and state, the Configuration
 *       copy, logging and partition-size accessors, and NeMo's exceptions
 *       (duplicate neuron, src/test/test.cpp:755-763; unknown neuron; the CPU
 *       backend, which is not part of this library)
 *   cpp_api_check run NETFILE RASTERFILE WEIGHTFILE
 *       a seeded 1000-neuron 80/20 Izhikevich network with 100 synapses per
 *       neuron, delays 1..20 and plastic excitatory synapses, built through
 *       nemo::Network; 2000 cycles of Simulation::step (firing stimulus of
 *       neurons 1 and 2 and current stimulus into neurons 3..5 in cycle 0)
 *       with applyStdp(1.0) after cycles 1000 and 2000. The network and the
 *       STDP window go to NETFILE, the fired ids of every cycle to RASTERFILE
 *       and every synapse weight read back with getSynapseWeight to
 *       WEIGHTFILE, for the test to compare with the oracle.
 *
 * Exit status 0 and "ok" on stdout on success; otherwise a message on stderr
 * and status 1.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "nemo.hpp"

static int failures = 0;

#define CHECK(cond) do { if(!(cond)) { \
	std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); failures++; } } while(0)

/* the status code of the nemo::exception the call throws, or -1 */
template<typename F>
static int thrown(F f)
{
	try {
		f();
	} catch(nemo::exception& e) {
		return e.errorNumber();
	}
	return -1;
}

static int network_checks()
{
	nemo::Network net;
	const unsigned iz = net.addNeuronType("Izhikevich");
	const float args[7] = {0.02f, 0.2f, -65.0f, 8.0f, 5.0f, -13.0f, -65.0f};
	net.addNeuron(iz, 10, 7, args);
	net.addNeuron(11, 0.02f, 0.2f, -65.0f, 8.0f, -13.0f, -65.0f, 0.0f);
	net.addNeuron(iz, 12, 7, args);
	CHECK(net.neuronCount() == 3);
	CHECK(net.maxDelay() == 0 && net.minWeight() == 0.0f && net.maxWeight() == 0.0f);
	const synapse_id a = net.addSynapse(10, 11, 7, 0.25f, 1);
	const synapse_id b = net.addSynapse(10, 12, 3, 0.5f, 0);
	const synapse_id c = net.addSynapse(11, 10, 13, -0.75f, 0);
	CHECK(b - a == 1u && (a >> 32) == 10u && (c >> 32) == 11u && (c & 0xffffffffu) == 0u);
	CHECK(net.maxDelay() == 13 && net.minWeight() == -0.75f && net.maxWeight() == 0.5f);
	CHECK(net.getSynapseSource(b) == 10 && net.getSynapseTarget(b) == 12 && net.getSynapseDelay(b) == 3);
	CHECK(net.getSynapseWeight(a) == 0.25f && net.getSynapsePlastic(a) == 1 && net.getSynapsePlastic(b) == 0);
	const std::vector<synapse_id>& from = net.getSynapsesFrom(10);
	CHECK(from.size() == 2 && from[0] == a && from[1] == b);
	/* parameters a,b,c,d,sigma then state u,v (src/nemo/plugins/Izhikevich.h) */
	CHECK(net.getNeuronParameter(10, 2) == -65.0f && net.getNeuronState(10, 1) == -65.0f);
	net.setNeuronState(12, 1, -70.0f);
	net.setNeuronParameter(12, 4, 2.0f);
	CHECK(net.getNeuronState(12, 1) == -70.0f && net.getNeuronParameter(12, 4) == 2.0f);
	/* adding an existing index throws (src/test/test.cpp:755-763) */
	CHECK(thrown([&] { net.addNeuron(iz, 10, 7, args); }) == NEMO_INVALID_INPUT);
	CHECK(thrown([&] { net.getNeuronState(99, 0); }) == NEMO_INVALID_INPUT);

	nemo::Configuration conf;
	CHECK(!conf.loggingEnabled());
	conf.enableLogging();
	conf.setCudaPartitionSize(256);
	conf.setWriteOnlySynapses();
	const nemo::Configuration copy(conf);
	conf.disableLogging();
	CHECK(!conf.loggingEnabled() && copy.loggingEnabled());
	CHECK(copy.cudaPartitionSize() == 256 && copy.writeOnlySynapses());
	CHECK(conf.backend() == NEMO_BACKEND_CUDA);
	/* NeMo's CPU backend is libnemo's own, not this library's */
	CHECK(thrown([&] { conf.setCpuBackend(); }) == NEMO_API_UNSUPPORTED);
	CHECK(conf.backend() == NEMO_BACKEND_CUDA);
	/* an STDP window longer than 64 cycles in all is refused (src/nemo/StdpFunction.cpp) */
	std::vector<float> longWindow(40, 0.1f);
	CHECK(thrown([&] { conf.setStdpFunction(longWindow, longWindow, -1.0f, 1.0f); }) != -1);
	CHECK(std::strlen(nemo::version()) > 0);
	return failures;
}

/* xorshift64* -> uniform float in [0, 1) and integers in [0, n) */
static uint64_t rs = 0x2545f4914f6cdd1dull;

static uint64_t next_u64()
{
	rs ^= rs >> 12;
	rs ^= rs << 25;
	rs ^= rs >> 27;
	return rs * 0x2545f4914f6cdd1dull;
}

static float uniform()
{
	return float(next_u64() >> 40) / float(1u << 24);
}

static int run(const char* netfile, const char* rasterfile, const char* weightfile)
{
	const unsigned ncount = 1000, scount = 100;
	std::FILE* nf = std::fopen(netfile, "wb");
	std::FILE* rf = std::fopen(rasterfile, "w");
	std::FILE* wf = std::fopen(weightfile, "wb");
	if(!nf || !rf || !wf) {
		std::fprintf(stderr, "cannot open the output files\n");
		return 1;
	}
	std::vector<float> args(7u * ncount), wgt(ncount * scount);
	std::vector<uint32_t> tgt(ncount * scount);
	std::vector<uint8_t> dly(ncount * scount), pls(ncount * scount);

	nemo::Network net;
	const unsigned iz = net.addNeuronType("Izhikevich");
	for(unsigned n = 0; n < ncount; ++n) {
		float* a = &args[7u * n];   /* a, b, c, d, sigma, u, v */
		const float v = -65.0f;
		const float r = uniform();
		if(n < (ncount * 4u) / 5u) {
			a[0] = 0.02f; a[1] = 0.2f; a[2] = v + 15.0f * r * r; a[3] = 8.0f - 6.0f * r * r; a[4] = 5.0f;
		} else {
			a[0] = 0.02f + 0.08f * r; a[1] = 0.25f - 0.05f * r; a[2] = v; a[3] = 2.0f; a[4] = 2.0f;
		}
		a[5] = a[1] * v;
		a[6] = v;
		net.addNeuron(iz, n, 7, a);
		for(unsigned s = 0; s < scount; ++s) {
			const unsigned k = n * scount + s;
			const bool exc = n < (ncount * 4u) / 5u;
			tgt[k] = uint32_t(next_u64() % ncount);
			dly[k] = uint8_t(1u + next_u64() % 20u);
			wgt[k] = exc ? 0.5f * uniform() : -uniform();
			pls[k] = exc ? 1 : 0;
			net.addSynapse(n, tgt[k], dly[k], wgt[k], pls[k]);
		}
	}
	/* exponential window, 20 samples each side */
	std::vector<float> pre(20), post(20);
	for(unsigned i = 0; i < 20; ++i) {
		pre[i] = 0.1f * std::exp(-float(i + 1) / 20.0f);
		post[i] = -0.12f * std::exp(-float(i + 1) / 20.0f);
	}
	const uint32_t hdr[2] = {ncount, scount};
	std::fwrite(hdr, sizeof(hdr), 1, nf);
	std::fwrite(args.data(), sizeof(float), args.size(), nf);
	std::fwrite(tgt.data(), sizeof(uint32_t), tgt.size(), nf);
	std::fwrite(dly.data(), 1, dly.size(), nf);
	std::fwrite(wgt.data(), sizeof(float), wgt.size(), nf);
	std::fwrite(pls.data(), 1, pls.size(), nf);
	std::fwrite(pre.data(), sizeof(float), pre.size(), nf);
	std::fwrite(post.data(), sizeof(float), post.size(), nf);
	std::fclose(nf);

	nemo::Configuration conf;
	conf.setStdpFunction(pre, post, -1.0f, 1.0f);
	std::unique_ptr<nemo::Simulation> sim(nemo::simulation(net, conf));
	if(!sim) {
		std::fprintf(stderr, "nemo::simulation: %s\n", nemo_strerror());
		return 1;
	}
	std::fprintf(stderr, "backend: %s\n", conf.backendDescription());
	nemo::Simulation::firing_stimulus fstim;
	fstim.push_back(1);
	fstim.push_back(2);
	nemo::Simulation::current_stimulus istim;
	for(unsigned n = 3; n < 6; ++n) istim.push_back(std::make_pair(n, 15.0f));
	for(unsigned t = 0; t < 2000; ++t) {
		const nemo::Simulation::firing_output& fired = (t == 0) ? sim->step(fstim, istim) : sim->step();
		std::fprintf(rf, "%u", t);
		for(unsigned x : fired) std::fprintf(rf, " %u", x);
		std::fprintf(rf, "\n");
		if(t % 1000 == 999) {
			sim->applyStdp(1.0f);
		}
	}
	std::fclose(rf);
	CHECK(sim->elapsedSimulation() == 2000);
	for(unsigned n = 0; n < ncount; ++n) {
		const std::vector<synapse_id>& ids = sim->getSynapsesFrom(n);
		CHECK(ids.size() == scount);
		for(size_t i = 0; i < ids.size(); ++i) {
			const float w = sim->getSynapseWeight(ids[i]);
			std::fwrite(&w, sizeof(float), 1, wf);
		}
	}
	std::fclose(wf);
	CHECK(sim->getSynapseTarget(synapse_id(5) << 32) == tgt[5 * scount]);
	const float v0 = sim->getMembranePotential(0);
	CHECK(v0 == sim->getNeuronState(0, 1));
	sim->setNeuronState(0, 1, -64.0f);
	CHECK(sim->getMembranePotential(0) == -64.0f);
	CHECK(thrown([&] { sim->getMembranePotential(ncount + 7); }) == NEMO_INVALID_INPUT);
	return failures;
}

int main(int argc, char** argv)
{
	int rc;
	try {
		if(argc == 2 && std::strcmp(argv[1], "net") == 0) {
			rc = network_checks();
		} else if(argc == 5 && std::strcmp(argv[1], "run") == 0) {
			rc = run(argv[2], argv[3], argv[4]);
		} else {
			std::fprintf(stderr, "usage: %s net | run NETFILE RASTERFILE WEIGHTFILE\n", argv[0]);
			return 2;
		}
	} catch(nemo::exception& e) {
		std::fprintf(stderr, "nemo::exception %d: %s\n", e.errorNumber(), e.what());
		return 1;
	}
	if(rc == 0) std::printf("ok\n");
	return rc ? 1 : 0;
}
