#include "common.hpp"

#include <cstdlib>
#include <deque>
#include <fstream>
#include <mutex>

#include <dlfcn.h>
#include <sys/stat.h>

#include <hip/hip_runtime.h>

namespace nemo {

/* built-in models and every descriptor loaded so far; entries never move */
static std::deque<ModelInfo>&
registry()
{
	static std::deque<ModelInfo> r;
	static bool init = false;
	if(!init) {
		init = true;
		/* name, id, nparam, nstate, membrane potential, normal RNG, builtin */
		r.push_back(ModelInfo{ "Izhikevich", 0, 5, 2, 1, true, true, "" });
		r.push_back(ModelInfo{ "IF_curr_exp", 1, 9, 4, 0, false, true, "" });
		r.push_back(ModelInfo{ "IF_curr_exp_noise", 2, 10, 4, 0, true, true, "" });
		r.push_back(ModelInfo{ "PoissonSource", 3, 1, 0, 0, true, true, "" });
		r.push_back(ModelInfo{ "Input", 4, 0, 0, 0, false, true, "" });
	}
	return r;
}

static std::mutex g_modelMutex;
static std::vector<std::string> g_userPaths;

static bool
isDirectory(const std::string& p)
{
	struct stat st;
	return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

static bool
isFile(const std::string& p)
{
	struct stat st;
	return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

void
addPluginPath(const std::string& dir)
{
	if(!isDirectory(dir)) {
		throw exception(NEMO_INVALID_INPUT, fmt("Plugin directory '", dir, "' does not exist"));
	}
	std::lock_guard<std::mutex> lock(g_modelMutex);
	g_userPaths.insert(g_userPaths.begin(), dir);
}

std::vector<std::string>
pluginPaths()
{
	/* src/nemo/NeuronType.cpp:50-67: paths added by the program, then the
	 * user directory ($HOME/.nemo/plugins), then the system directory (the
	 * plugins/ directory beside this library) */
	std::vector<std::string> p;
	{
		std::lock_guard<std::mutex> lock(g_modelMutex);
		p = g_userPaths;
	}
	const char* home = getenv("HOME");
	if(home && *home) {
		const std::string user = std::string(home) + "/.nemo/plugins";
		if(isDirectory(user)) p.push_back(user);
	}
	Dl_info info;
	if(dladdr(reinterpret_cast<void*>(&pluginPaths), &info) && info.dli_fname) {
		std::string lib(info.dli_fname);
		size_t slash = lib.rfind('/');
		std::string libdir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
		p.push_back(libdir + "/plugins");
	}
	return p;
}

static bool
isTrue(const std::string& v)
{
	return v == "true" || v == "1" || v == "yes" || v == "on";
}

/* Minimal reader of NeMo's model descriptor format: key=value lines,
 * [section] headers, '#' comments (src/nemo/NeuronType.cpp:92-155). */
static bool
readDescriptor(const std::string& path, ModelInfo& m)
{
	std::ifstream f(path.c_str());
	if(!f) return false;
	std::string line, section;
	bool haveP = false, haveS = false;
	m.membranePotential = 0;
	m.normalRng = false;
	while(std::getline(f, line)) {
		size_t h = line.find('#');
		if(h != std::string::npos) line = line.substr(0, h);
		size_t a = line.find_first_not_of(" \t\r");
		if(a == std::string::npos) continue;
		line = line.substr(a);
		size_t b = line.find_last_not_of(" \t\r");
		line = line.substr(0, b + 1);
		if(line[0] == '[') {
			section = line.substr(1, line.find(']') - 1);
			continue;
		}
		size_t eq = line.find('=');
		if(eq == std::string::npos) continue;
		std::string key = line.substr(0, eq), val = line.substr(eq + 1);
		while(!key.empty() && (key.back() == ' ' || key.back() == '\t')) key.pop_back();
		size_t v0 = val.find_first_not_of(" \t");
		val = v0 == std::string::npos ? std::string() : val.substr(v0);
		if(section.empty() && key == "parameters") { m.nparam = unsigned(std::atoi(val.c_str())); haveP = true; }
		else if(section.empty() && key == "state-variables") { m.nstate = unsigned(std::atoi(val.c_str())); haveS = true; }
		else if(section.empty() && key == "membrane-potential") m.membranePotential = unsigned(std::atoi(val.c_str()));
		else if(section.empty() && key == "history") m.history = unsigned(std::atoi(val.c_str()));
		else if(section == "rng" && key == "normal") m.normalRng = isTrue(val);
		else if(section == "rcm" && key == "sources") m.rcmSources = isTrue(val);
		else if(section == "rcm" && key == "delays") m.rcmDelays = isTrue(val);
		else if(section == "rcm" && key == "forward") m.rcmForward = isTrue(val);
		else if(section == "rcm" && key == "weights") m.rcmWeights = isTrue(val);
	}
	if(m.history < 1) {
		throw exception(NEMO_DL_ERROR, fmt("Model descriptor ", path, " has history=0; at least one cycle is kept"));
	}
	if(!haveP || !haveS) {
		throw exception(NEMO_DL_ERROR, fmt("Model descriptor ", path, " lacks parameters= or state-variables="));
	}
	if(m.nparam + m.nstate > 16) {
		/* NEMO_MAX_NEURON_ARGS, src/nemo/config.h.in:32 */
		throw exception(NEMO_DL_ERROR, fmt("Model descriptor ", path, " has more than 16 parameters and state variables"));
	}
	return true;
}

const ModelInfo*
findModel(const std::string& name)
{
	{
		std::lock_guard<std::mutex> lock(g_modelMutex);
		for(const ModelInfo& m : registry()) {
			if(m.name == name) return &m;
		}
	}
	for(const std::string& dir : pluginPaths()) {
		std::string ini = dir + "/" + name + ".ini";
		if(!isFile(ini)) continue;
		ModelInfo m;
		m.name = name;
		m.builtin = false;
		m.dir = dir;
		if(readDescriptor(ini, m)) {
			std::lock_guard<std::mutex> lock(g_modelMutex);
			m.id = unsigned(registry().size());
			registry().push_back(m);
			return &registry().back();
		}
	}
	throw exception(NEMO_DL_ERROR, fmt("Unknown neuron model '", name,
				"': no built-in model and no ", name, ".ini on the plugin path"));
}

const ModelInfo&
modelInfo(unsigned id)
{
	std::lock_guard<std::mutex> lock(g_modelMutex);
	return registry().at(id);
}

int32_t
fx_toFix(float f)
{
	if(std::abs(int(f)) >= 1 << (31 - 20)) {
		throw exception(NEMO_INVALID_INPUT, fmt("Fixed-point overflow. Value ", f,
					" does not fit into fixed-point format Q11.20"));
	}
	return static_cast<int32_t>(f * (1 << 20));
}

float
fx_toFloat(int32_t v)
{
	return float(v) / float(1 << 20);
}

void
hipCheck(int err, const char* what)
{
	if(err != hipSuccess) {
		int code = (err == hipErrorOutOfMemory) ? NEMO_CUDA_MEMORY_ERROR : NEMO_CUDA_INVOCATION_ERROR;
		throw exception(code, fmt(what, ": ", hipGetErrorString(hipError_t(err))));
	}
}

void
DeviceMutex::lock()
{
	std::unique_lock<std::mutex> g(m_);
	const std::thread::id me = std::this_thread::get_id();
	if(m_depth > 0 && m_owner == me) {
		m_depth += 1;
		return;
	}
	const uint64_t ticket = m_next++;
	cv_.wait(g, [&] { return m_depth == 0 && m_serving == ticket; });
	m_owner = me;
	m_depth = 1;
}

void
DeviceMutex::unlock()
{
	std::lock_guard<std::mutex> g(m_);
	if(--m_depth == 0) {
		m_owner = std::thread::id();
		m_serving += 1;
		cv_.notify_all();
	}
}

DeviceMutex&
deviceMutex()
{
	/* never destroyed: threads may still hold it while the process exits */
	static DeviceMutex* m = new DeviceMutex();
	return *m;
}

} // namespace nemo
