/*
 * c_api_check.c -- NeMo's C-API tests (src/test/c_api.cpp) as a plain C99
 * program built against include/nemo.h and linked with libnemo_hip.so, the
 * way a NeMo C user would use the library. tests/test_c_api.py runs it.
 *
 *   c_api_check ids
 *       synapse ids are (source << 32) | ordinal (src/test/c_api.cpp:336-355);
 *       network only, no device needed
 *   c_api_check synapses N0
 *       a triangular network (neuron N0+n has synapses to N0..N0+n-1, delay
 *       1 + s%20, weight s%10, plastic s%2) answers every synapse query the
 *       same from the network and from the simulation (src/test/c_api.cpp:
 *       524-597)
 *   c_api_check run NETFILE RASTERFILE FSTIM ISTIM
 *       builds a seeded 1000 x 1000 80/20 Izhikevich network of delay-1
 *       synapses through the C API, runs 2000 cycles with nemo_step (firing
 *       stimulus of neuron 100 and/or current stimulus 20 into neurons 20,
 *       40, 60 in cycle 0 when FSTIM / ISTIM are 1; NULL arrays otherwise),
 *       reading synapses and membrane potentials every 100 cycles, and then
 *       replaces neuron 0 (src/test/c_api.cpp:155-315). The network goes to
 *       NETFILE and the fired ids of every cycle to RASTERFILE, for the test
 *       to compare with the oracle.
 *
 * Exit status 0 on success; on failure a message on stderr and status 1.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nemo.h"

static int failures = 0;

#define CHECK(cond) do { if(!(cond)) { \
	fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); failures++; } } while(0)

static void call(nemo_status_t s, const char* what)
{
	if(s != NEMO_OK) {
		fprintf(stderr, "%s failed (status %d): %s\n", what, (int) s, nemo_strerror());
		exit(1);
	}
}

static int synapse_ids(void)
{
	nemo_network_t net = nemo_new_network();
	synapse_id id00, id01, id10;
	call(nemo_add_synapse(net, 0, 1, 1, 0.0f, 0, &id00), "nemo_add_synapse");
	call(nemo_add_synapse(net, 0, 1, 1, 0.0f, 0, &id01), "nemo_add_synapse");
	call(nemo_add_synapse(net, 1, 0, 1, 0.0f, 0, &id10), "nemo_add_synapse");
	CHECK(id01 - id00 == 1u);
	CHECK((id10 & 0xffffffffu) == 0u);
	CHECK((id00 & 0xffffffffu) == 0u);
	CHECK((id01 & 0xffffffffu) == 1u);
	CHECK((id00 >> 32) == 0u);
	CHECK((id01 >> 32) == 0u);
	CHECK((id10 >> 32) == 1u);
	nemo_delete_network(net);
	return failures;
}

/* Q11.20 round trip of a weight, as the simulation stores it */
static float fx20(float w)
{
	return (float) (int32_t) (w * (float) (1 << 20)) / (float) (1 << 20);
}

static int get_synapses(unsigned n0)
{
	const unsigned ncount = 1000;
	nemo_network_t net = nemo_new_network();
	for(unsigned n = 0; n < ncount; ++n) {
		call(nemo_add_neuron_iz(net, n0 + n, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f), "nemo_add_neuron_iz");
		for(unsigned s = 0; s < n; ++s) {
			synapse_id id;
			call(nemo_add_synapse(net, n0 + n, n0 + s, 1 + s % 20, (float) (s % 10), (unsigned char) (s % 2), &id),
					"nemo_add_synapse");
		}
	}
	nemo_configuration_t conf = nemo_new_configuration();
	nemo_simulation_t sim = nemo_new_simulation(net, conf);
	if(sim == NULL) {
		fprintf(stderr, "nemo_new_simulation: %s\n", nemo_strerror());
		exit(1);
	}
	for(unsigned src = n0; src < n0 + ncount; ++src) {
		synapse_id* ids;
		size_t nlen, slen;
		call(nemo_get_synapses_from_n(net, src, &ids, &nlen), "nemo_get_synapses_from_n");
		call(nemo_get_synapses_from_s(sim, src, &ids, &slen), "nemo_get_synapses_from_s");
		CHECK(nlen == slen);
		CHECK(slen == src - n0);
		for(unsigned i = 0; i < slen; ++i) {
			const synapse_id id = ids[i];
			unsigned ns, ss, nt, st, nd, sd;
			float nw, sw;
			unsigned char np, sp;
			call(nemo_get_synapse_source_n(net, id, &ns), "source_n");
			call(nemo_get_synapse_source_s(sim, id, &ss), "source_s");
			call(nemo_get_synapse_target_n(net, id, &nt), "target_n");
			call(nemo_get_synapse_target_s(sim, id, &st), "target_s");
			call(nemo_get_synapse_delay_n(net, id, &nd), "delay_n");
			call(nemo_get_synapse_delay_s(sim, id, &sd), "delay_s");
			call(nemo_get_synapse_weight_n(net, id, &nw), "weight_n");
			call(nemo_get_synapse_weight_s(sim, id, &sw), "weight_s");
			call(nemo_get_synapse_plastic_n(net, id, &np), "plastic_n");
			call(nemo_get_synapse_plastic_s(sim, id, &sp), "plastic_s");
			CHECK(ns == ss && ns == src);
			CHECK(nt == st && nt == n0 + i);
			CHECK(nd == sd && nd == 1 + i % 20);
			CHECK(fx20(nw) == sw && nw == (float) (i % 10));
			CHECK(np == sp && np == i % 2);
			if(failures) {
				fprintf(stderr, "at source %u synapse %u\n", src, i);
				return failures;
			}
		}
	}
	nemo_delete_simulation(sim);
	nemo_delete_configuration(conf);
	nemo_delete_network(net);
	return failures;
}

/* xorshift64* -> uniform float in [0, 1) and integers in [0, n) */
static uint64_t rs = 0x9e3779b97f4a7c15ull;

static uint64_t next_u64(void)
{
	rs ^= rs >> 12;
	rs ^= rs << 25;
	rs ^= rs >> 27;
	return rs * 0x2545f4914f6cdd1dull;
}

static float uniform(void)
{
	return (float) (next_u64() >> 40) / (float) (1u << 24);
}

static int run(const char* netfile, const char* rasterfile, int use_fstim, int use_istim)
{
	const unsigned ncount = 1000, scount = 1000;
	FILE* nf = fopen(netfile, "wb");
	FILE* rf = fopen(rasterfile, "w");
	if(nf == NULL || rf == NULL) {
		fprintf(stderr, "cannot open the output files\n");
		return 1;
	}
	float* args = malloc(sizeof(float) * 7u * ncount);
	uint32_t* tgt = malloc(sizeof(uint32_t) * ncount * scount);
	float* wgt = malloc(sizeof(float) * ncount * scount);
	if(args == NULL || tgt == NULL || wgt == NULL) {
		fprintf(stderr, "out of host memory\n");
		return 1;
	}

	nemo_network_t net = nemo_new_network();
	unsigned iz;
	call(nemo_add_neuron_type(net, "Izhikevich", &iz), "nemo_add_neuron_type");
	for(unsigned n = 0; n < ncount; ++n) {
		float* a = args + 7u * n;   /* a, b, c, d, sigma, u, v */
		const float v = -65.0f;
		const float r1 = uniform(), r2 = uniform();
		if(n < (ncount * 4u) / 5u) {
			a[0] = 0.02f; a[1] = 0.2f; a[2] = v + 15.0f * r1 * r1; a[3] = 8.0f - 6.0f * r2 * r2; a[4] = 5.0f;
		} else {
			a[0] = 0.02f + 0.08f * r1; a[1] = 0.25f - 0.05f * r2; a[2] = v; a[3] = 2.0f; a[4] = 2.0f;
		}
		a[5] = a[1] * v;
		a[6] = v;
		call(nemo_add_neuron(net, iz, n, 7, a), "nemo_add_neuron");
		for(unsigned s = 0; s < scount; ++s) {
			const unsigned k = n * scount + s;
			tgt[k] = (uint32_t) (next_u64() % ncount);
			wgt[k] = (n < (ncount * 4u) / 5u) ? 0.5f * uniform() : -uniform();
			call(nemo_add_synapse(net, n, tgt[k], 1, wgt[k], 0, NULL), "nemo_add_synapse");
		}
	}
	const uint32_t hdr[2] = {ncount, scount};
	fwrite(hdr, sizeof(hdr), 1, nf);
	fwrite(args, sizeof(float), 7u * ncount, nf);
	fwrite(tgt, sizeof(uint32_t), ncount * scount, nf);
	fwrite(wgt, sizeof(float), ncount * scount, nf);
	fclose(nf);

	unsigned fstim[1] = {100};
	unsigned inidx[3] = {20, 40, 60};
	float icur[3] = {20.0f, 20.0f, 20.0f};

	nemo_configuration_t conf = nemo_new_configuration();
	const char* descr;
	call(nemo_backend_description(conf, &descr), "nemo_backend_description");
	fprintf(stderr, "backend: %s\n", descr);
	nemo_simulation_t sim = nemo_new_simulation(net, conf);
	if(sim == NULL) {
		fprintf(stderr, "nemo_new_simulation: %s\n", nemo_strerror());
		return 1;
	}
	for(unsigned t = 0; t < 2000; ++t) {
		unsigned* fired;
		size_t nfired;
		if(t == 0) {
			call(nemo_step(sim, use_fstim ? fstim : NULL, use_fstim ? 1 : 0,
						use_istim ? inidx : NULL, use_istim ? icur : NULL, use_istim ? 3 : 0,
						&fired, &nfired), "nemo_step");
		} else {
			call(nemo_step(sim, NULL, 0, NULL, NULL, 0, &fired, &nfired), "nemo_step");
		}
		fprintf(rf, "%u", t);
		for(size_t i = 0; i < nfired; ++i) fprintf(rf, " %u", fired[i]);
		fprintf(rf, "\n");
		if(t % 100 == 0) {
			synapse_id* ids;
			size_t len;
			float w, v;
			unsigned target, delay;
			unsigned char plastic;
			call(nemo_get_synapses_from_s(sim, 1, &ids, &len), "nemo_get_synapses_from_s");
			CHECK(len == scount);
			call(nemo_get_synapse_weight_s(sim, ids[0], &w), "nemo_get_synapse_weight_s");
			call(nemo_get_synapse_target_s(sim, ids[0], &target), "nemo_get_synapse_target_s");
			call(nemo_get_synapse_delay_s(sim, ids[0], &delay), "nemo_get_synapse_delay_s");
			call(nemo_get_synapse_plastic_s(sim, ids[0], &plastic), "nemo_get_synapse_plastic_s");
			CHECK(target == tgt[scount] && delay == 1 && plastic == 0 && w == fx20(wgt[scount]));
			call(nemo_get_membrane_potential(sim, 40, &v), "nemo_get_membrane_potential");
			call(nemo_get_membrane_potential(sim, 50, &v), "nemo_get_membrane_potential");
		}
	}
	fclose(rf);
	{
		float v;
		call(nemo_set_neuron_iz_s(sim, 0, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -64.0f, 0.0f), "nemo_set_neuron_iz_s");
		call(nemo_get_membrane_potential(sim, 0, &v), "nemo_get_membrane_potential");
		CHECK(v == -64.0f);
	}
	nemo_delete_simulation(sim);
	nemo_delete_configuration(conf);
	nemo_delete_network(net);
	free(args);
	free(tgt);
	free(wgt);
	return failures;
}

int main(int argc, char** argv)
{
	int rc;
	if(argc == 2 && strcmp(argv[1], "ids") == 0) {
		rc = synapse_ids();
	} else if(argc == 3 && strcmp(argv[1], "synapses") == 0) {
		rc = get_synapses((unsigned) strtoul(argv[2], NULL, 10));
	} else if(argc == 6 && strcmp(argv[1], "run") == 0) {
		rc = run(argv[2], argv[3], atoi(argv[4]), atoi(argv[5]));
	} else {
		fprintf(stderr, "usage: %s ids | synapses N0 | run NETFILE RASTERFILE FSTIM ISTIM\n", argv[0]);
		return 2;
	}
	if(rc == 0) printf("ok\n");
	return rc ? 1 : 0;
}
