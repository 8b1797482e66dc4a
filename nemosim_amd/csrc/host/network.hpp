#ifndef NEMOSIM_HOST_NETWORK_HPP
#define NEMOSIM_HOST_NETWORK_HPP

#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "common.hpp"

namespace nemo {
namespace hip {

/* Neurons of one type, structure of arrays in insertion order
 * (src/nemo/Neurons.cpp:29-47). */
struct NeuronGroup {
	const ModelInfo* model;
	std::vector<std::vector<float> > param;   /* [nparam][n] */
	std::vector<std::vector<float> > state;   /* [nstate][n] */
	std::vector<unsigned> gidx;
};

/*
 * Network under construction (NeMo's NetworkImpl, src/nemo/NetworkImpl.cpp).
 * Synapses are kept in flat arrays in insertion order; synapse ids are
 * (source << 32) | (index among the source's synapses), as
 * src/nemo/synapse_indices.hpp:30-51.
 */
class NetworkImpl
{
	public:
		NetworkImpl();

		unsigned addNeuronType(const std::string& name);
		void addNeuron(unsigned type, unsigned gidx, unsigned nargs, const float args[]);
		void addNeurons(unsigned type, size_t n, const unsigned idx[], unsigned nargs, const float args[]);
		void setNeuron(unsigned gidx, unsigned nargs, const float args[]);
		float getNeuronState(unsigned gidx, unsigned var) const;
		float getNeuronParameter(unsigned gidx, unsigned param) const;
		void setNeuronState(unsigned gidx, unsigned var, float val);
		void setNeuronParameter(unsigned gidx, unsigned param, float val);

		synapse_id addSynapse(unsigned source, unsigned target, unsigned delay,
				float weight, unsigned char plastic);
		void addSynapsesCsr(size_t nsources, const unsigned sources[], const uint64_t rowptr[],
				const unsigned targets[], const unsigned char delays[],
				const float weights[], const unsigned char plastic[]);

		unsigned getSynapseSource(synapse_id id) const;
		unsigned getSynapseTarget(synapse_id id) const;
		unsigned getSynapseDelay(synapse_id id) const;
		float getSynapseWeight(synapse_id id) const;
		unsigned char getSynapsePlastic(synapse_id id) const;
		const std::vector<synapse_id>& getSynapsesFrom(unsigned source);

		unsigned neuronCount() const { return unsigned(m_mapper.size()); }
		unsigned maxDelay() const { return m_maxDelay; }
		size_t synapseCount() const { return m_tgt.size(); }

		/* storage read by the simulation builder */
		std::vector<NeuronGroup> groups;
		const std::vector<unsigned>& src() const { return m_src; }
		const std::vector<unsigned>& tgt() const { return m_tgt; }
		const std::vector<unsigned char>& delay() const { return m_delay; }
		const std::vector<float>& weight() const { return m_weight; }
		const std::vector<unsigned char>& plastic() const { return m_plastic; }

		bool hasNeuron(unsigned gidx) const { return m_mapper.count(gidx) != 0; }

	private:
		std::map<std::string, unsigned> m_typeIds;
		std::unordered_map<unsigned, std::pair<unsigned, unsigned> > m_mapper;

		std::vector<unsigned> m_src, m_tgt;
		std::vector<unsigned char> m_delay, m_plastic;
		std::vector<float> m_weight;
		std::unordered_map<unsigned, uint32_t> m_outCount;
		unsigned m_maxDelay;

		/* lazily built: flat index of each synapse by (source, ordinal) */
		mutable bool m_indexValid;
		mutable std::unordered_map<unsigned, std::pair<size_t, size_t> > m_rows;
		mutable std::vector<size_t> m_order;
		std::vector<synapse_id> m_queried;

		const std::pair<unsigned, unsigned>& address(unsigned gidx) const;
		size_t flatIndex(synapse_id id) const;
		void buildIndex() const;
		void appendSynapse(unsigned s, unsigned t, unsigned d, float w, unsigned char p);
};

} // namespace hip
} // namespace nemo

#endif
