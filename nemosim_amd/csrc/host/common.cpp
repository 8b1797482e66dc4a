#include "common.hpp"

#include <cstdlib>
#include <hip/hip_runtime.h>

namespace nemo {

static const ModelInfo MODELS[] = {
	/* name, id, nparam, nstate, membrane potential, normal RNG */
	{ "Izhikevich", 0, 5, 2, 1, true },
	{ "IF_curr_exp", 1, 9, 4, 0, false },
	{ "IF_curr_exp_noise", 2, 10, 4, 0, true },
};

const ModelInfo*
findModel(const std::string& name)
{
	for(const ModelInfo& m : MODELS) {
		if(name == m.name) return &m;
	}
	return nullptr;
}

const ModelInfo&
modelInfo(unsigned id)
{
	return MODELS[id];
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

} // namespace nemo
