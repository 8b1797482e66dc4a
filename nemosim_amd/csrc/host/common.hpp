#ifndef NEMOSIM_HOST_COMMON_HPP
#define NEMOSIM_HOST_COMMON_HPP

#include <condition_variable>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>
#include <sstream>
#include <vector>
#include <stdint.h>

#include "nemo_types.h"

namespace nemo {

/* NeMo's exception with a status code, src/nemo/exception.hpp:23-37 */
class exception : public std::runtime_error
{
	public:
		exception(int errorNumber, const std::string& msg) :
			std::runtime_error(msg), m_errno(errorNumber) {}
		int errorNumber() const { return m_errno; }
	private:
		int m_errno;
};

template<typename... Args>
std::string fmt(Args&&... args)
{
	std::ostringstream s;
	using expand = int[];
	(void) expand{0, ((s << args), 0)...};
	return s.str();
}

/* A neuron model: its descriptor (NeMo's <Model>.ini: parameters=,
 * state-variables=, membrane-potential=, history=, [rcm] sources= delays=
 * forward= weights=, [rng] normal=, src/nemo/NeuronType.cpp:92-155) and
 * where its plugin lives. Built-in models (Izhikevich, IF_curr_exp,
 * IF_curr_exp_noise) also run in the fused step kernel; any other model
 * found on the plugin path runs through its plugin's hip_update_neurons. */
struct ModelInfo {
	std::string name;
	unsigned id;
	unsigned nparam;
	unsigned nstate;
	unsigned membranePotential;
	bool normalRng;
	bool builtin;
	std::string dir;      /* directory of the descriptor; plugin in dir/hip/ */
	unsigned history = 1;          /* cycles of state kept */
	bool rcmSources = false;       /* reverse-matrix fields the model reads */
	bool rcmDelays = false;
	bool rcmForward = false;
	bool rcmWeights = false;
	bool usesRcm() const { return rcmSources || rcmDelays || rcmForward || rcmWeights; }
};

/* Resolve a model name: built-ins first, then <dir>/<name>.ini on the plugin
 * path (most recently added directory first, then the library's own plugin
 * directory). Throws NEMO_DL_ERROR if not found. */
const ModelInfo* findModel(const std::string& name);
const ModelInfo& modelInfo(unsigned id);

/* The plugin search path (src/nemo/Plugin.cpp:84-121): user directories,
 * most recent first, then <directory of libnemo_hip.so>/plugins. */
void addPluginPath(const std::string& dir);
std::vector<std::string> pluginPaths();

/* Q11.20 conversion, src/nemo/fixedpoint.cpp:18-29 (truncating; throws on
 * overflow) */
int32_t fx_toFix(float f);
float fx_toFloat(int32_t v);

void hipCheck(int err, const char* what);

/* A recursive mutex that is granted in the order it was asked for: a
 * thread that releases it and asks again queues behind the threads already
 * waiting, so a loop of launches on one thread cannot starve uploads queued
 * by others. */
class DeviceMutex
{
	public:
		void lock();
		void unlock();

	private:
		std::mutex m_;
		std::condition_variable cv_;
		uint64_t m_next = 0;        /* next ticket to hand out */
		uint64_t m_serving = 0;     /* ticket whose holder may take it */
		unsigned m_depth = 0;
		std::thread::id m_owner;
};

/* Held around this library's device work -- uploads at construction,
 * launches, copies -- so that work queued from one host thread (a serving
 * loop building simulations on worker threads) never shares the device
 * with a persistent step launch: its workgroups must all be resident at
 * once, and a copy kernel dispatched into the middle of that launch could
 * delay some of them. Recursive: a batched run launches single runs. First
 * come, first served: a serving loop that launches chunk after chunk lets
 * the uploads of the networks built meanwhile in between. */
DeviceMutex& deviceMutex();

} // namespace nemo

#endif
