#ifndef NEMO_H
#define NEMO_H

/*
 * nemo.h -- C API of the nemosim HIP backend for the NeMo spiking neural
 * network simulator (NeMo 0.7.2 surface).
 *
 * Every declaration below keeps the name, argument order, argument meaning
 * and status-code behaviour of the NeMo function it replaces; the comment on
 * each cites that function in NeMo's src/nemo.h. A program written against
 * NeMo's C API links against libnemo_hip.so unchanged. Simulations created
 * here run on an AMD Instinct GPU (gfx950); see nemo_hip.h for the HIP
 * specific extensions (bulk construction, multi-step runs, batching).
 *
 * Errors: every call that can fail returns a nemo_status_t (NEMO_OK == 0);
 * the message of the last failure is returned by nemo_strerror(), a single
 * process-wide string (as in NeMo, so calls are not thread safe).
 */

#include <stddef.h>
#include "nemo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque handles (src/nemo.h:42-50) */
struct nemo_network_class;
struct nemo_simulation_class;
struct nemo_configuration_class;
typedef struct nemo_network_class* nemo_network_t;
typedef struct nemo_simulation_class* nemo_simulation_t;
typedef struct nemo_configuration_class* nemo_configuration_t;

typedef int nemo_status_t;

/* src/nemo.h:56-57 */
const char* nemo_version(void);

/* src/nemo.h:72-74: prepend a directory to the neuron-model plugin path */
nemo_status_t nemo_add_plugin_path(const char* dir);

/* ---------------------------------------------------------------- hardware */

/* src/nemo.h:88-90 (the CUDA device query; here: HIP devices) */
nemo_status_t nemo_cuda_device_count(unsigned* count);

/* src/nemo.h:93-95 */
nemo_status_t nemo_cuda_device_description(unsigned device, const char** descr);

/* ----------------------------------------------------------- configuration */

/* src/nemo.h:105-106 */
nemo_configuration_t nemo_new_configuration(void);

/* src/nemo.h:109-110 */
void nemo_delete_configuration(nemo_configuration_t);

/* src/nemo.h:114-115 */
nemo_status_t nemo_log_stdout(nemo_configuration_t);

/* src/nemo.h:148-156 */
nemo_status_t nemo_set_stdp_function(nemo_configuration_t,
		float prefire_fn[], size_t prefire_len,
		float postfire_fn[], size_t postfire_len,
		float min_excitatory_weight,
		float max_excitatory_weight,
		float min_inhibitory_weight,
		float max_inhibitory_weight);

/* src/nemo.h:160-162. The CPU backend is NeMo's own (libnemo_cpu); selecting
 * it here returns NEMO_API_UNSUPPORTED. */
nemo_status_t nemo_set_cpu_backend(nemo_configuration_t);

/* src/nemo.h:166-168: the GPU backend slot; dev < 0 picks device 0 */
nemo_status_t nemo_set_cuda_backend(nemo_configuration_t conf, int dev);

/* src/nemo.h:171-173 */
nemo_status_t nemo_cuda_device(nemo_configuration_t conf, int* dev);

/* src/nemo.h:177-179 */
nemo_status_t nemo_backend(nemo_configuration_t conf, backend_t* backend);

/* src/nemo.h:183-185 */
nemo_status_t nemo_backend_description(nemo_configuration_t conf, const char** descr);

/* src/nemo.h:189-191 */
nemo_status_t nemo_set_write_only_synapses(nemo_configuration_t conf);

/* ------------------------------------------------------------ construction */

/* src/nemo.h:215-216 */
nemo_network_t nemo_new_network(void);

/* src/nemo.h:220-221 */
void nemo_delete_network(nemo_network_t);

/* src/nemo.h:235-239 */
nemo_status_t nemo_add_neuron_type(nemo_network_t, const char* name, unsigned* neuron_type);

/* src/nemo.h:245-250 (deprecated Izhikevich form; args a,b,c,d,u,v,sigma) */
nemo_status_t nemo_add_neuron_iz(nemo_network_t, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);

/* src/nemo.h:265-268 */
nemo_status_t nemo_add_neuron(nemo_network_t, unsigned type, unsigned idx,
		unsigned nargs, float args[]);

/* src/nemo.h:278-286 */
nemo_status_t nemo_add_synapse(nemo_network_t,
		unsigned source, unsigned target, unsigned delay, float weight,
		unsigned char is_plastic, synapse_id* id);

/* src/nemo.h:290-292 */
nemo_status_t nemo_neuron_count(nemo_network_t net, unsigned* ncount);

/* -------------------------------------------------------------- simulation */

/* src/nemo.h:311-312. Returns NULL on failure (message in nemo_strerror). */
nemo_simulation_t nemo_new_simulation(nemo_network_t, nemo_configuration_t);

/* src/nemo.h:316-317 */
void nemo_delete_simulation(nemo_simulation_t);

/* src/nemo.h:347-352. *fired points into backend storage valid until the
 * next step of the same simulation. */
nemo_status_t nemo_step(nemo_simulation_t,
		unsigned fstim_nidx[], size_t fstim_count,
		unsigned istim_nidx[], float istim_current[], size_t istim_count,
		unsigned* fired[], size_t* fired_count);

/* src/nemo.h:356-358 */
nemo_status_t nemo_apply_stdp(nemo_simulation_t, float reward);

/* ----------------------------------------------------------------- queries */

/* src/nemo.h:380-382 */
nemo_status_t nemo_get_membrane_potential(nemo_simulation_t sim, unsigned neuron, float* v);

/* src/nemo.h:399-401 */
nemo_status_t nemo_get_neuron_state_n(nemo_network_t net, unsigned neuron, unsigned var, float* val);

/* src/nemo.h:418-420 */
nemo_status_t nemo_get_neuron_parameter_n(nemo_network_t net, unsigned neuron, unsigned param, float* val);

/* src/nemo.h:437-439 */
nemo_status_t nemo_get_neuron_state_s(nemo_simulation_t sim, unsigned neuron, unsigned var, float* val);

/* src/nemo.h:456-458 */
nemo_status_t nemo_get_neuron_parameter_s(nemo_simulation_t sim, unsigned neuron, unsigned param, float* val);

/* src/nemo.h:467-469 */
nemo_status_t nemo_get_synapse_source_n(nemo_network_t, synapse_id synapse, unsigned* source);

/* src/nemo.h:478-480 */
nemo_status_t nemo_get_synapse_source_s(nemo_simulation_t, synapse_id synapse, unsigned* source);

/* src/nemo.h:489-491 */
nemo_status_t nemo_get_synapse_target_n(nemo_network_t, synapse_id synapse, unsigned* target);

/* src/nemo.h:499-501 */
nemo_status_t nemo_get_synapse_target_s(nemo_simulation_t, synapse_id synapse, unsigned* target);

/* src/nemo.h:509-511 */
nemo_status_t nemo_get_synapse_delay_n(nemo_network_t, synapse_id synapse, unsigned* delay);

/* src/nemo.h:519-521 */
nemo_status_t nemo_get_synapse_delay_s(nemo_simulation_t, synapse_id synapse, unsigned* delay);

/* src/nemo.h:529-531 */
nemo_status_t nemo_get_synapse_weight_n(nemo_network_t, synapse_id synapse, float* weight);

/* src/nemo.h:539-541 */
nemo_status_t nemo_get_synapse_weight_s(nemo_simulation_t, synapse_id synapse, float* weight);

/* src/nemo.h:549-551 */
nemo_status_t nemo_get_synapse_plastic_n(nemo_network_t, synapse_id synapse, unsigned char* plastic);

/* src/nemo.h:559-561 */
nemo_status_t nemo_get_synapse_plastic_s(nemo_simulation_t, synapse_id synapse, unsigned char* plastic);

/* src/nemo.h:565-567 */
nemo_status_t nemo_get_synapses_from_n(nemo_network_t, unsigned source, synapse_id* synapses[], size_t* len);

/* src/nemo.h:579-581 */
nemo_status_t nemo_get_synapses_from_s(nemo_simulation_t, unsigned source, synapse_id* synapses[], size_t* len);

/* ------------------------------------------------------------ modification */

/* src/nemo.h:613-615 */
nemo_status_t nemo_set_neuron_n(nemo_network_t, unsigned idx, unsigned nargs, float args[]);

/* src/nemo.h:627-629 */
nemo_status_t nemo_set_neuron_s(nemo_simulation_t, unsigned idx, unsigned nargs, float args[]);

/* src/nemo.h:639-644 */
nemo_status_t nemo_set_neuron_iz_n(nemo_network_t net, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);

/* src/nemo.h:654-659 */
nemo_status_t nemo_set_neuron_iz_s(nemo_simulation_t sim, unsigned idx,
		float a, float b, float c, float d, float u, float v, float sigma);

/* src/nemo.h:675-677 */
nemo_status_t nemo_set_neuron_state_n(nemo_network_t net, unsigned neuron, unsigned var, float val);

/* src/nemo.h:693-695 */
nemo_status_t nemo_set_neuron_parameter_n(nemo_network_t net, unsigned neuron, unsigned param, float val);

/* src/nemo.h:711-713 */
nemo_status_t nemo_set_neuron_state_s(nemo_simulation_t sim, unsigned neuron, unsigned var, float val);

/* src/nemo.h:729-731 */
nemo_status_t nemo_set_neuron_parameter_s(nemo_simulation_t sim, unsigned neuron, unsigned param, float val);

/* ------------------------------------------------------------------ timers */

/* src/nemo.h:752-754 */
nemo_status_t nemo_elapsed_wallclock(nemo_simulation_t, unsigned long*);

/* src/nemo.h:756-758 */
nemo_status_t nemo_elapsed_simulation(nemo_simulation_t, unsigned long*);

/* src/nemo.h:760-762 */
nemo_status_t nemo_reset_timer(nemo_simulation_t);

/* ------------------------------------------------------------------ errors */

/* src/nemo.h:786-787 */
const char* nemo_strerror(void);

#ifdef __cplusplus
}
#endif

#endif
