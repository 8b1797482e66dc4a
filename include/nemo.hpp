#ifndef NEMO_HPP
#define NEMO_HPP

/*
 * nemo.hpp -- NeMo's C++ API (src/nemo.hpp, src/nemo/Network.hpp,
 * src/nemo/Configuration.hpp, src/nemo/Simulation.hpp) over the HIP
 * backend's C ABI (nemo.h, nemo_hip.h). Header-only: a program written
 * against NeMo's C++ classes compiles against this header and links
 * against libnemo_hip.so; no C++ object crosses the library boundary.
 *
 * Errors are thrown as nemo::exception carrying NeMo's status code
 * (src/nemo/exception.hpp:23-37), with the message of nemo_strerror().
 * As in NeMo, the classes are not copy-assignable and a Simulation is used
 * from one host thread at a time.
 */

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nemo.h"
#include "nemo_hip.h"

namespace nemo {

/* src/nemo/exception.hpp:23-37 */
class exception : public std::runtime_error
{
	public:
		exception(int errorNumber, const std::string& msg) :
			std::runtime_error(msg), m_errno(errorNumber) {}
		int errorNumber() const { return m_errno; }
	private:
		int m_errno;
};

namespace detail {

inline void check(nemo_status_t s)
{
	if(s != NEMO_OK) {
		throw nemo::exception(s, nemo_strerror());
	}
}

} // namespace detail

class Simulation;
class Network;
class Configuration;
Simulation* simulation(const Network& net, const Configuration& conf);

/* src/nemo/Configuration.hpp:44-121 */
class Configuration
{
	public:
		Configuration() : m_h(nemo_new_configuration())
		{
			if(!m_h) throw nemo::exception(NEMO_ALLOCATION_ERROR, "configuration allocation failed");
		}

		Configuration(const Configuration& other) : m_h(nemo_copy_configuration(other.m_h))
		{
			if(!m_h) detail::check(nemo_hip_last_status());
		}

		~Configuration() { nemo_delete_configuration(m_h); }

		void enableLogging() { detail::check(nemo_log_stdout(m_h)); }
		void disableLogging() { detail::check(nemo_disable_logging(m_h)); }
		bool loggingEnabled() const
		{
			int on = 0;
			detail::check(nemo_logging_enabled(m_h, &on));
			return on != 0;
		}

		void setCudaPartitionSize(unsigned ps) { detail::check(nemo_set_partition_size(m_h, ps)); }
		unsigned cudaPartitionSize() const
		{
			unsigned ps = 0;
			detail::check(nemo_partition_size(m_h, &ps));
			return ps;
		}

		/* one weight range for both signs (src/nemo/Configuration.cpp:79-87):
		 * excitatory [0, maxWeight], inhibitory [minWeight, 0] */
		void setStdpFunction(const std::vector<float>& prefire, const std::vector<float>& postfire,
				float minWeight, float maxWeight)
		{
			setStdpFunction(prefire, postfire, 0.0f, maxWeight, 0.0f, minWeight);
		}

		void setStdpFunction(const std::vector<float>& prefire, const std::vector<float>& postfire,
				float minExcitatoryWeight, float maxExcitatoryWeight,
				float minInhibitoryWeight, float maxInhibitoryWeight)
		{
			std::vector<float> pre(prefire), post(postfire);
			detail::check(nemo_set_stdp_function(m_h, pre.empty() ? nullptr : &pre[0], pre.size(),
						post.empty() ? nullptr : &post[0], post.size(),
						minExcitatoryWeight, maxExcitatoryWeight, minInhibitoryWeight, maxInhibitoryWeight));
		}

		void setWriteOnlySynapses() { detail::check(nemo_set_write_only_synapses(m_h)); }
		bool writeOnlySynapses() const
		{
			int on = 0;
			detail::check(nemo_write_only_synapses(m_h, &on));
			return on != 0;
		}

		/* the HIP device takes NeMo's GPU backend slot */
		void setCudaBackend(int device = -1) { detail::check(nemo_set_cuda_backend(m_h, device)); }
		/* NeMo's CPU backend is not part of this library: throws
		 * NEMO_API_UNSUPPORTED */
		void setCpuBackend() { detail::check(nemo_set_cpu_backend(m_h)); }

		backend_t backend() const
		{
			backend_t b;
			detail::check(nemo_backend(m_h, &b));
			return b;
		}

		int cudaDevice() const
		{
			int d = -1;
			detail::check(nemo_cuda_device(m_h, &d));
			return d;
		}

		const char* backendDescription() const
		{
			const char* s = nullptr;
			detail::check(nemo_backend_description(m_h, &s));
			return s;
		}

		nemo_configuration_t handle() const { return m_h; }

	private:
		nemo_configuration_t m_h;
		Configuration& operator=(const Configuration&) = delete;
};

/* src/nemo/Network.hpp:38-190 and the queries of src/nemo/ReadableNetwork.hpp */
class Network
{
	public:
		Network() : m_h(nemo_new_network())
		{
			if(!m_h) throw nemo::exception(NEMO_ALLOCATION_ERROR, "network allocation failed");
		}

		~Network() { nemo_delete_network(m_h); }

		unsigned addNeuronType(const std::string& name)
		{
			unsigned t = 0;
			detail::check(nemo_add_neuron_type(m_h, name.c_str(), &t));
			return t;
		}

		void addNeuron(unsigned type, unsigned idx, unsigned nargs, const float args[])
		{
			std::vector<float> a(args, args + nargs);
			detail::check(nemo_add_neuron(m_h, type, idx, nargs, a.empty() ? nullptr : &a[0]));
		}

		/* Izhikevich neuron (deprecated in NeMo in favour of the generic form) */
		void addNeuron(unsigned idx, float a, float b, float c, float d, float u, float v, float sigma)
		{
			detail::check(nemo_add_neuron_iz(m_h, idx, a, b, c, d, u, v, sigma));
		}

		void setNeuron(unsigned idx, unsigned nargs, const float args[])
		{
			std::vector<float> a(args, args + nargs);
			detail::check(nemo_set_neuron_n(m_h, idx, nargs, a.empty() ? nullptr : &a[0]));
		}

		void setNeuron(unsigned idx, float a, float b, float c, float d, float u, float v, float sigma)
		{
			detail::check(nemo_set_neuron_iz_n(m_h, idx, a, b, c, d, u, v, sigma));
		}

		synapse_id addSynapse(unsigned source, unsigned target, unsigned delay, float weight, unsigned char plastic)
		{
			synapse_id id = 0;
			detail::check(nemo_add_synapse(m_h, source, target, delay, weight, plastic, &id));
			return id;
		}

		float getNeuronParameter(unsigned neuron, unsigned parameter) const
		{
			float x = 0.0f;
			detail::check(nemo_get_neuron_parameter_n(m_h, neuron, parameter, &x));
			return x;
		}

		float getNeuronState(unsigned neuron, unsigned var) const
		{
			float x = 0.0f;
			detail::check(nemo_get_neuron_state_n(m_h, neuron, var, &x));
			return x;
		}

		void setNeuronParameter(unsigned neuron, unsigned parameter, float value)
		{
			detail::check(nemo_set_neuron_parameter_n(m_h, neuron, parameter, value));
		}

		void setNeuronState(unsigned neuron, unsigned var, float value)
		{
			detail::check(nemo_set_neuron_state_n(m_h, neuron, var, value));
		}

		unsigned getSynapseSource(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_source_n(m_h, id, &x));
			return x;
		}

		unsigned getSynapseTarget(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_target_n(m_h, id, &x));
			return x;
		}

		unsigned getSynapseDelay(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_delay_n(m_h, id, &x));
			return x;
		}

		float getSynapseWeight(const synapse_id& id) const
		{
			float x = 0.0f;
			detail::check(nemo_get_synapse_weight_n(m_h, id, &x));
			return x;
		}

		unsigned char getSynapsePlastic(const synapse_id& id) const
		{
			unsigned char x = 0;
			detail::check(nemo_get_synapse_plastic_n(m_h, id, &x));
			return x;
		}

		const std::vector<synapse_id>& getSynapsesFrom(unsigned neuron)
		{
			synapse_id* ids = nullptr;
			size_t len = 0;
			detail::check(nemo_get_synapses_from_n(m_h, neuron, &ids, &len));
			m_synapses.assign(ids, ids + len);
			return m_synapses;
		}

		unsigned maxDelay() const
		{
			unsigned d = 0;
			detail::check(nemo_max_delay(m_h, &d));
			return d;
		}

		float maxWeight() const
		{
			float w = 0.0f;
			detail::check(nemo_max_weight(m_h, &w));
			return w;
		}

		float minWeight() const
		{
			float w = 0.0f;
			detail::check(nemo_min_weight(m_h, &w));
			return w;
		}

		unsigned neuronCount() const
		{
			unsigned n = 0;
			detail::check(nemo_neuron_count(m_h, &n));
			return n;
		}

		nemo_network_t handle() const { return m_h; }

	private:
		nemo_network_t m_h;
		std::vector<synapse_id> m_synapses;
		Network(const Network&) = delete;
		Network& operator=(const Network&) = delete;
};

/* src/nemo/Simulation.hpp:38-218. Created by nemo::simulation(); the caller
 * deletes it. */
class Simulation
{
	public:
		typedef std::vector<unsigned> firing_output;
		typedef std::vector<unsigned> firing_stimulus;
		typedef std::vector< std::pair<unsigned, float> > current_stimulus;

		virtual ~Simulation() { nemo_delete_simulation(m_h); }

		/* one 1 ms cycle; the result is valid until the next step */
		const firing_output& step() { return step(firing_stimulus(), current_stimulus()); }
		const firing_output& step(const firing_stimulus& fstim) { return step(fstim, current_stimulus()); }
		const firing_output& step(const current_stimulus& istim) { return step(firing_stimulus(), istim); }

		const firing_output& step(const firing_stimulus& fstim, const current_stimulus& istim)
		{
			std::vector<unsigned> fs(fstim), ii;
			std::vector<float> iv;
			ii.reserve(istim.size());
			iv.reserve(istim.size());
			for(size_t i = 0; i < istim.size(); ++i) {
				ii.push_back(istim[i].first);
				iv.push_back(istim[i].second);
			}
			unsigned* fired = nullptr;
			size_t nfired = 0;
			detail::check(nemo_step(m_h, fs.empty() ? nullptr : &fs[0], fs.size(),
						ii.empty() ? nullptr : &ii[0], iv.empty() ? nullptr : &iv[0], ii.size(),
						&fired, &nfired));
			m_fired.assign(fired, fired + nfired);
			return m_fired;
		}

		void setNeuron(unsigned idx, unsigned nargs, const float args[])
		{
			std::vector<float> a(args, args + nargs);
			detail::check(nemo_set_neuron_s(m_h, idx, nargs, a.empty() ? nullptr : &a[0]));
		}

		void setNeuronState(unsigned neuron, unsigned var, float value)
		{
			detail::check(nemo_set_neuron_state_s(m_h, neuron, var, value));
		}

		void setNeuronParameter(unsigned neuron, unsigned parameter, float value)
		{
			detail::check(nemo_set_neuron_parameter_s(m_h, neuron, parameter, value));
		}

		void applyStdp(float reward) { detail::check(nemo_apply_stdp(m_h, reward)); }

		float getNeuronState(unsigned neuron, unsigned var) const
		{
			float x = 0.0f;
			detail::check(nemo_get_neuron_state_s(m_h, neuron, var, &x));
			return x;
		}

		float getNeuronParameter(unsigned neuron, unsigned parameter) const
		{
			float x = 0.0f;
			detail::check(nemo_get_neuron_parameter_s(m_h, neuron, parameter, &x));
			return x;
		}

		float getMembranePotential(unsigned neuron) const
		{
			float v = 0.0f;
			detail::check(nemo_get_membrane_potential(m_h, neuron, &v));
			return v;
		}

		const std::vector<synapse_id>& getSynapsesFrom(unsigned neuron)
		{
			synapse_id* ids = nullptr;
			size_t len = 0;
			detail::check(nemo_get_synapses_from_s(m_h, neuron, &ids, &len));
			m_synapses.assign(ids, ids + len);
			return m_synapses;
		}

		unsigned getSynapseSource(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_source_s(m_h, id, &x));
			return x;
		}

		unsigned getSynapseTarget(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_target_s(m_h, id, &x));
			return x;
		}

		unsigned getSynapseDelay(const synapse_id& id) const
		{
			unsigned x = 0;
			detail::check(nemo_get_synapse_delay_s(m_h, id, &x));
			return x;
		}

		float getSynapseWeight(const synapse_id& id) const
		{
			float x = 0.0f;
			detail::check(nemo_get_synapse_weight_s(m_h, id, &x));
			return x;
		}

		unsigned char getSynapsePlastic(const synapse_id& id) const
		{
			unsigned char x = 0;
			detail::check(nemo_get_synapse_plastic_s(m_h, id, &x));
			return x;
		}

		unsigned long elapsedWallclock() const
		{
			unsigned long ms = 0;
			detail::check(nemo_elapsed_wallclock(m_h, &ms));
			return ms;
		}

		unsigned long elapsedSimulation() const
		{
			unsigned long ms = 0;
			detail::check(nemo_elapsed_simulation(m_h, &ms));
			return ms;
		}

		void resetTimer() { detail::check(nemo_reset_timer(m_h)); }

		nemo_simulation_t handle() const { return m_h; }

	private:
		friend Simulation* simulation(const Network& net, const Configuration& conf);
		explicit Simulation(nemo_simulation_t h) : m_h(h) {}
		nemo_simulation_t m_h;
		firing_output m_fired;
		std::vector<synapse_id> m_synapses;
		Simulation(const Simulation&) = delete;
		Simulation& operator=(const Simulation&) = delete;
};

/* src/nemo.hpp:34-36: a simulation of net on the configured device. NeMo
 * returns NULL when no backend can take it; the reason is then in
 * nemo_strerror(). */
inline Simulation* simulation(const Network& net, const Configuration& conf)
{
	nemo_simulation_t h = nemo_new_simulation(net.handle(), conf.handle());
	if(!h) return nullptr;
	return new Simulation(h);
}

/* src/nemo.hpp:39-52 */
inline unsigned cudaDeviceCount()
{
	unsigned n = 0;
	detail::check(nemo_cuda_device_count(&n));
	return n;
}

inline const char* cudaDeviceDescription(unsigned device)
{
	const char* s = nullptr;
	detail::check(nemo_cuda_device_description(device, &s));
	return s;
}

/* src/nemo.hpp:55-58 */
inline const char* version() { return nemo_version(); }

/* src/nemo.hpp:62-65 */
inline void addPluginPath(const std::string& dir) { detail::check(nemo_add_plugin_path(dir.c_str())); }

} // namespace nemo

#endif
