/*
 * batch.hpp -- many independent networks advanced together: packed into
 * co-resident launches of the fused step kernel, each launch replayed from
 * an instantiated HIP graph (nemo_hip_run_batch).
 */
#ifndef NEMOSIM_BATCH_HPP
#define NEMOSIM_BATCH_HPP

#include <cstddef>

namespace nemo {
namespace hip {

class Simulation;

/* Runs nsteps[i] cycles of sims[i] for every i. Networks of one device and
 * step-kernel variant are packed greedily, in the given order, while their
 * workgroups fit co-resident; simulations on the plugin path run on their own. events[i] (if events) receives the
 * delivered-synapse count of sims[i]. */
void runBatch(Simulation* const sims[], size_t n, const unsigned nsteps[], unsigned long long events[]);

} // namespace hip
} // namespace nemo

#endif
