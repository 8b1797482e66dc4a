#include "network.hpp"

#include <algorithm>

namespace nemo {
namespace hip {

NetworkImpl::NetworkImpl() :
	m_maxDelay(0),
	m_indexValid(false)
{
}

/* src/nemo/NetworkImpl.cpp:38-48: a name maps to one type id. Model names
 * resolve to a built-in model or a descriptor on the plugin path
 * (src/nemo/NeuronType.cpp:46-70). */
unsigned
NetworkImpl::addNeuronType(const std::string& name)
{
	std::map<std::string, unsigned>::const_iterator i = m_typeIds.find(name);
	if(i != m_typeIds.end()) {
		return i->second;
	}
	const ModelInfo* m = findModel(name);
	unsigned id = unsigned(groups.size());
	NeuronGroup g;
	g.model = m;
	g.param.resize(m->nparam);
	g.state.resize(m->nstate);
	groups.push_back(g);
	m_typeIds[name] = id;
	return id;
}

/* src/nemo/NetworkImpl.cpp:84-92 and src/nemo/Neurons.cpp:29-47. The
 * duplicate-index check runs before the neuron is stored (NeMo stores it
 * first and throws from the mapper; either way the call fails). */
void
NetworkImpl::addNeuron(unsigned type, unsigned gidx, unsigned nargs, const float args[])
{
	if(type >= groups.size()) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid neuron type id ", type));
	}
	NeuronGroup& g = groups[type];
	const unsigned np = g.model->nparam, ns = g.model->nstate;
	if(nargs != np + ns) {
		throw exception(NEMO_INVALID_INPUT, fmt("Unexpected number of parameters/state variables when adding neuron. Expected ",
					np + ns, ", found ", nargs));
	}
	if(m_mapper.count(gidx)) {
		throw exception(NEMO_INVALID_INPUT, fmt("Duplicate neuron index ", gidx));
	}
	for(unsigned i = 0; i < np; ++i) g.param[i].push_back(args[i]);
	for(unsigned i = 0; i < ns; ++i) g.state[i].push_back(args[np + i]);
	m_mapper[gidx] = std::make_pair(type, unsigned(g.gidx.size()));
	g.gidx.push_back(gidx);
}

void
NetworkImpl::addNeurons(unsigned type, size_t n, const unsigned idx[], unsigned nargs, const float args[])
{
	if(type < groups.size()) {
		NeuronGroup& g = groups[type];
		for(auto& v : g.param) v.reserve(v.size() + n);
		for(auto& v : g.state) v.reserve(v.size() + n);
		g.gidx.reserve(g.gidx.size() + n);
		m_mapper.reserve(m_mapper.size() + n);
	}
	for(size_t i = 0; i < n; ++i) {
		addNeuron(type, idx[i], nargs, args + i * nargs);
	}
}

const std::pair<unsigned, unsigned>&
NetworkImpl::address(unsigned gidx) const
{
	auto i = m_mapper.find(gidx);
	if(i == m_mapper.end()) {
		throw exception(NEMO_INVALID_INPUT, fmt("Non-existing neuron index ", gidx));
	}
	return i->second;
}

void
NetworkImpl::setNeuron(unsigned gidx, unsigned nargs, const float args[])
{
	const auto& a = address(gidx);
	NeuronGroup& g = groups[a.first];
	const unsigned np = g.model->nparam, ns = g.model->nstate;
	if(nargs != np + ns) {
		throw exception(NEMO_INVALID_INPUT, fmt("Unexpected number of parameters/state variables when modifying neuron. Expected ",
					np + ns, ", found ", nargs));
	}
	for(unsigned i = 0; i < np; ++i) g.param[i][a.second] = args[i];
	for(unsigned i = 0; i < ns; ++i) g.state[i][a.second] = args[np + i];
}

float
NetworkImpl::getNeuronState(unsigned gidx, unsigned var) const
{
	const auto& a = address(gidx);
	const NeuronGroup& g = groups[a.first];
	if(var >= g.model->nstate) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid state variable index ", var));
	}
	return g.state[var][a.second];
}

float
NetworkImpl::getNeuronParameter(unsigned gidx, unsigned param) const
{
	const auto& a = address(gidx);
	const NeuronGroup& g = groups[a.first];
	if(param >= g.model->nparam) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid parameter index ", param));
	}
	return g.param[param][a.second];
}

void
NetworkImpl::setNeuronState(unsigned gidx, unsigned var, float val)
{
	const auto& a = address(gidx);
	NeuronGroup& g = groups[a.first];
	if(var >= g.model->nstate) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid state variable index ", var));
	}
	g.state[var][a.second] = val;
}

void
NetworkImpl::setNeuronParameter(unsigned gidx, unsigned param, float val)
{
	const auto& a = address(gidx);
	NeuronGroup& g = groups[a.first];
	if(param >= g.model->nparam) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid parameter index ", param));
	}
	g.param[param][a.second] = val;
}

void
NetworkImpl::appendSynapse(unsigned s, unsigned t, unsigned d, float w, unsigned char p)
{
	/* src/nemo/NetworkImpl.cpp:103-108 */
	if(d < 1) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid delay (", d, ") for synapse between ", s, " and ", t));
	}
	if(d > 255) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid delay (", d, ") for synapse between ", s, " and ", t,
					"; the HIP backend supports delays up to 64"));
	}
	m_src.push_back(s);
	m_tgt.push_back(t);
	m_delay.push_back((unsigned char) d);
	m_weight.push_back(w);
	m_plastic.push_back(p ? 1 : 0);
	m_maxDelay = std::max(m_maxDelay, d);
	m_indexValid = false;
}

synapse_id
NetworkImpl::addSynapse(unsigned source, unsigned target, unsigned delay, float weight, unsigned char plastic)
{
	appendSynapse(source, target, delay, weight, plastic);
	uint32_t& c = m_outCount[source];
	synapse_id id = (synapse_id(source) << 32) | synapse_id(c);
	c += 1;
	return id;
}

void
NetworkImpl::addSynapsesCsr(size_t nsources, const unsigned sources[], const uint64_t rowptr[],
		const unsigned targets[], const unsigned char delays[],
		const float weights[], const unsigned char plastic[])
{
	size_t total = size_t(rowptr[nsources] - rowptr[0]);
	m_src.reserve(m_src.size() + total);
	m_tgt.reserve(m_tgt.size() + total);
	m_delay.reserve(m_delay.size() + total);
	m_weight.reserve(m_weight.size() + total);
	m_plastic.reserve(m_plastic.size() + total);
	for(size_t k = 0; k < nsources; ++k) {
		unsigned s = sources ? sources[k] : unsigned(k);
		uint64_t b = rowptr[k], e = rowptr[k + 1];
		for(uint64_t i = b; i < e; ++i) {
			appendSynapse(s, targets[i], delays[i], weights[i], plastic ? plastic[i] : 0);
		}
		if(e > b) {
			m_outCount[s] += uint32_t(e - b);
		}
	}
}

void
NetworkImpl::buildIndex() const
{
	if(m_indexValid) return;
	/* stable counting sort of the flat synapse indices by source */
	m_rows.clear();
	std::unordered_map<unsigned, size_t> count;
	for(unsigned s : m_src) count[s] += 1;
	size_t base = 0;
	for(auto& c : count) {
		m_rows[c.first] = std::make_pair(base, size_t(0));
		base += c.second;
	}
	m_order.assign(m_src.size(), 0);
	for(size_t i = 0; i < m_src.size(); ++i) {
		auto& r = m_rows[m_src[i]];
		m_order[r.first + r.second] = i;
		r.second += 1;
	}
	m_indexValid = true;
}

size_t
NetworkImpl::flatIndex(synapse_id id) const
{
	buildIndex();
	unsigned source = unsigned(id >> 32);
	size_t k = size_t(id & 0xffffffffu);
	auto r = m_rows.find(source);
	if(r == m_rows.end()) {
		throw exception(NEMO_INVALID_INPUT, fmt("synapses of non-existing neuron (", source, ") requested"));
	}
	if(k >= r->second.second) {
		throw exception(NEMO_INVALID_INPUT, "synapse id out or range");
	}
	return m_order[r->second.first + k];
}

unsigned NetworkImpl::getSynapseSource(synapse_id id) const { return m_src[flatIndex(id)]; }
unsigned NetworkImpl::getSynapseTarget(synapse_id id) const { return m_tgt[flatIndex(id)]; }
unsigned NetworkImpl::getSynapseDelay(synapse_id id) const { return m_delay[flatIndex(id)]; }
float NetworkImpl::getSynapseWeight(synapse_id id) const { return m_weight[flatIndex(id)]; }
unsigned char NetworkImpl::getSynapsePlastic(synapse_id id) const { return m_plastic[flatIndex(id)]; }

/* src/nemo/NetworkImpl.cpp:216-230 */
const std::vector<synapse_id>&
NetworkImpl::getSynapsesFrom(unsigned source)
{
	buildIndex();
	auto r = m_rows.find(source);
	if(r == m_rows.end()) {
		if(!hasNeuron(source)) {
			throw exception(NEMO_INVALID_INPUT, fmt("synapses of non-existing neuron (", source, ") requested"));
		}
		m_queried.clear();
		return m_queried;
	}
	m_queried.resize(r->second.second);
	for(size_t k = 0; k < m_queried.size(); ++k) {
		m_queried[k] = (synapse_id(source) << 32) | synapse_id(k);
	}
	return m_queried;
}

} // namespace hip
} // namespace nemo
