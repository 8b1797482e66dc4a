/*
 * batch.cpp -- the batched form of the step (nemo_hip_run_batch).
 *
 * A pack of networks becomes one persistent launch whose work table maps
 * each workgroup to (network, partition), XCD-aware over the whole pack.
 * The launch is a HIP graph of three kernels -- clear the networks' status
 * words and event counters, the step kernel, gather status and events into
 * one array -- instantiated once per work table and replayed: every
 * per-launch value (cycle, step count, stimulus, flags) lives in the
 * device-side network table, rewritten before each replay with one copy
 * from pinned host memory. A replay therefore costs one copy in, one graph
 * launch, one copy out and one synchronisation, whatever the pack size.
 */
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "simulation.hpp"

namespace nemo {
namespace hip {

namespace {

struct BatchPlan {
	int device = 0;
	uint32_t variant = 0;
	uint32_t blockDim = 0;
	uint32_t lds = 0;
	uint32_t nnets = 0;
	std::vector<WorkItem> work;
	DevBuf dnets, dwork, dcollect;
	NetDev* hnets = nullptr;       /* pinned */
	uint32_t* hcollect = nullptr;  /* pinned, [nnets][16] */
	void* exec = nullptr;

	bool matches(int dev, uint32_t var, uint32_t bd, uint32_t ld, uint32_t nn, const std::vector<WorkItem>& w) const
	{
		if(dev != device || var != variant || bd != blockDim || ld != lds || nn != nnets
				|| w.size() != work.size()) return false;
		for(size_t i = 0; i < w.size(); ++i) {
			if(w[i].net != work[i].net || w[i].part != work[i].part) return false;
		}
		return true;
	}

	~BatchPlan()
	{
		(void) nemo_dev_graph_destroy(exec);
		if(hnets) (void) hipHostFree(hnets);
		if(hcollect) (void) hipHostFree(hcollect);
	}
};

/* Instantiated plans, most recently used first. Kept for the life of the
 * process (never destroyed at exit, when the runtime may be gone). */
std::mutex g_planMutex;
std::list<std::unique_ptr<BatchPlan> >* g_plans = new std::list<std::unique_ptr<BatchPlan> >();
const size_t MAX_PLANS = 32;
hipStream_t g_streams[64] = {};

hipStream_t
batchStream(int dev)
{
	if(dev < 0 || dev >= 64) throw exception(NEMO_INVALID_INPUT, "device index out of range");
	if(!g_streams[dev]) {
		hipCheck(hipStreamCreate(&g_streams[dev]), "hipStreamCreate");
	}
	return g_streams[dev];
}

/* the device wall clock's rate (wall_clock64 in the step kernel), kHz */
double
wallClockKHz(int dev)
{
	static double g_khz[64] = {};
	if(g_khz[dev] == 0.0) {
		int khz = 0;
		hipCheck(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
		g_khz[dev] = khz > 0 ? double(khz) : 100000.0;
	}
	return g_khz[dev];
}

BatchPlan&
getPlan(int dev, uint32_t variant, uint32_t blockDim, uint32_t lds, uint32_t nnets,
		const std::vector<WorkItem>& work)
{
	for(auto it = g_plans->begin(); it != g_plans->end(); ++it) {
		if((*it)->matches(dev, variant, blockDim, lds, nnets, work)) {
			g_plans->splice(g_plans->begin(), *g_plans, it);
			return *g_plans->front();
		}
	}
	std::unique_ptr<BatchPlan> plan(new BatchPlan());
	plan->device = dev;
	plan->variant = variant;
	plan->blockDim = blockDim;
	plan->lds = lds;
	plan->nnets = nnets;
	plan->work = work;
	plan->dnets.alloc(size_t(nnets) * sizeof(NetDev));
	plan->dwork.upload(work.data(), work.size() * sizeof(WorkItem));
	plan->dcollect.alloc(size_t(nnets) * 64);
	hipCheck(hipHostMalloc(reinterpret_cast<void**>(&plan->hnets), size_t(nnets) * sizeof(NetDev)), "hipHostMalloc");
	hipCheck(hipHostMalloc(reinterpret_cast<void**>(&plan->hcollect), size_t(nnets) * 64), "hipHostMalloc");
	hipCheck(nemo_dev_batch_graph_create(plan->dnets.as<NetDev>(), nnets, plan->dwork.as<WorkItem>(),
				uint32_t(work.size()), blockDim, lds, variant, plan->dcollect.as<uint32_t>(), &plan->exec),
			"batched step graph");
	g_plans->push_front(std::move(plan));
	while(g_plans->size() > MAX_PLANS) {
		g_plans->pop_back();
	}
	return *g_plans->front();
}

} // namespace

void
runBatch(Simulation* const sims[], size_t n, const unsigned nsteps[], unsigned long long events[])
{
	std::lock_guard<std::mutex> lock(g_planMutex);
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	/* Plugin-path simulations and empty work run on their own; the rest are
	 * grouped by (device, kernel variant) and, within a group, by workgroup
	 * size, keeping the given order within each size, then packed greedily
	 * into launches. A launch runs every workgroup at its largest member's
	 * size and LDS, so networks of like partitions share a launch and more
	 * of them fit; a batch that changes by a few networks between calls
	 * (serving) keeps most of its launches as they were. */
	std::vector<size_t> order;
	for(size_t i = 0; i < n; ++i) {
		Simulation* sim = sims[i];
		const unsigned ns = nsteps ? nsteps[i] : 0;
		if(!sim->fused() || sim->partitions() == 0 || ns == 0) {
			unsigned long long ev = 0;
			sim->run(ns, nullptr, &ev);
			if(events) events[i] = ev;
		} else {
			order.push_back(i);
		}
	}
	std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
		const Simulation* a = sims[x];
		const Simulation* b = sims[y];
		if(a->device() != b->device()) return a->device() < b->device();
		if(a->variant() != b->variant()) return a->variant() < b->variant();
		return a->blockThreads() < b->blockThreads();
	});
	size_t i = 0;
	while(i < order.size()) {
		Simulation* first = sims[order[i]];
		const int dev = first->device();
		const uint32_t variant = first->variant();
		hipCheck(hipSetDevice(dev), "hipSetDevice");
		uint32_t blockDim = 64, lds = 0, blocks = 0;
		size_t j = i;
		std::vector<size_t> members;
		while(j < order.size()) {
			Simulation* sim = sims[order[j]];
			if(sim->device() != dev || sim->variant() != variant) break;
			const uint32_t bd = std::max(blockDim, sim->blockThreads());
			const uint32_t ld = std::max(lds, sim->batchLdsBytes());
			uint32_t maxB = 0;
			hipCheck(nemo_dev_step_max_blocks(bd, ld, variant, &maxB), "occupancy query");
			if(blocks + sim->partitions() > maxB) {
				if(members.empty()) {
					throw exception(NEMO_API_UNSUPPORTED, "simulation does not fit the device");
				}
				break;
			}
			blockDim = bd;
			lds = ld;
			blocks += sim->partitions();
			members.push_back(order[j]);
			++j;
		}
		hipStream_t st = batchStream(dev);
		std::vector<NetDev> nets;
		std::vector<std::pair<uint32_t, uint32_t> > items;
		for(size_t k : members) {
			Simulation* sim = sims[k];
			const uint32_t idx = uint32_t(nets.size());
			nets.push_back(sim->prepareLaunch(nsteps[k], true));
			for(uint32_t p = 0; p < sim->partitions(); ++p) items.push_back(std::make_pair(idx, p));
		}
		const std::vector<WorkItem> work = Simulation::xcdWorkTable(items);
		BatchPlan& plan = getPlan(dev, variant, blockDim, lds, uint32_t(nets.size()), work);
		std::memcpy(plan.hnets, nets.data(), nets.size() * sizeof(NetDev));
		hipCheck(hipMemcpyAsync(plan.dnets.ptr, plan.hnets, nets.size() * sizeof(NetDev),
					hipMemcpyHostToDevice, st), "H2D");
		hipCheck(nemo_dev_graph_launch(plan.exec, st), "batched step graph launch");
		hipCheck(hipMemcpyAsync(plan.hcollect, plan.dcollect.ptr, nets.size() * 64,
					hipMemcpyDeviceToHost, st), "D2H");
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
		/* every member ran: advance all of them, then report the first error */
		bool failed = false;
		exception firstError(NEMO_OK, "");
		for(size_t m = 0; m < members.size(); ++m) {
			const size_t k = members[m];
			const uint32_t* c = plan.hcollect + 16 * m;
			if(events) {
				events[k] = (unsigned long long) c[2] | ((unsigned long long) c[3] << 32);
			}
			/* its workgroups' first start to last end on the device clock */
			const uint64_t t0 = (uint64_t) c[8] | ((uint64_t) c[9] << 32);
			const uint64_t t1 = (uint64_t) c[10] | ((uint64_t) c[11] << 32);
			sims[k]->setKernelMs(t1 > t0 ? float(double(t1 - t0) / wallClockKHz(dev)) : 0.0f);
			try {
				/* the graph cleared the block before the launch: c counts
				 * from zero, and the block keeps these values afterwards */
				sims[k]->setControlBase(c);
				sims[k]->completeLaunch(nsteps[k], c);
			} catch(exception& e) {
				if(!failed) {
					firstError = exception(e.errorNumber(), fmt(e.what(), " [batched launch of ", members.size(),
								" networks, ", work.size(), " workgroups of ", blockDim, " threads, ", lds,
								" B LDS, variant ", variant, "]"));
				}
				failed = true;
			}
		}
		if(failed) throw firstError;
		i = j;
	}
}

} // namespace hip
} // namespace nemo
