#ifndef NEMOSIM_HOST_COMMON_HPP
#define NEMOSIM_HOST_COMMON_HPP

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

/* Built-in neuron models. Counts and membrane-potential index follow the
 * model descriptors src/nemo/plugins/Izhikevich.ini:1-5 and
 * src/nemo/plugins/IF_curr_exp.ini:14-26; IF_curr_exp_noise is
 * IF_curr_exp with a 10th parameter sigma (Gaussian input current). */
struct ModelInfo {
	const char* name;
	unsigned id;
	unsigned nparam;
	unsigned nstate;
	unsigned membranePotential;
	bool normalRng;
};

const ModelInfo* findModel(const std::string& name);
const ModelInfo& modelInfo(unsigned id);

/* Q11.20 conversion, src/nemo/fixedpoint.cpp:18-29 (truncating; throws on
 * overflow) */
int32_t fx_toFix(float f);
float fx_toFloat(int32_t v);

void hipCheck(int err, const char* what);

} // namespace nemo

#endif
