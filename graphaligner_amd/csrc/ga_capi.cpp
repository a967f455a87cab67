// extern "C" boundary (include/graphaligner_amd.h).
#include "../../include/graphaligner_amd.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <chrono>
#include <thread>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <sys/mman.h>
#include "ga_align.hpp"
#include "ga_capi_internal.hpp"

using namespace ga;

namespace ga {

void copyErr(char* err, size_t cap, const char* msg)
{
	if (!err || cap == 0) return;
	std::strncpy(err, msg, cap - 1);
	err[cap - 1] = 0;
}

int32_t errorCostOf(const ga_params_t& p)
{
	// Params::XscoreErrorCost (src/GraphAlignerCommon.h:94)
	return (int32_t)(100 * (p.preciseClippingIdentity / (1.0 - p.preciseClippingIdentity) + 1.0));
}

int32_t xdropOf(const ga_params_t& p)
{
	if (p.xdrop >= 0) return p.xdrop;
	// src/AlignerMain.cpp:220-231
	double x = 50;
	if (p.preciseClippingIdentity >= 0.501) x = std::max(x, 50.0 * (100 * (p.preciseClippingIdentity / (1.0 - p.preciseClippingIdentity) + 1.0)));
	return (int32_t)x;
}

gad::ExtParams extParamsOf(const ga_params_t& p)
{
	gad::ExtParams e;
	e.bandwidth = (int32_t)p.bandwidth;
	e.errorCost = errorCostOf(p);
	e.xdrop = xdropOf(p);
	e.pad = 0;
	e.maxCellsPerSlice = p.maxCellsPerSlice <= 0 ? INT64_MAX : p.maxCellsPerSlice;
	return e;
}

} // namespace ga

extern "C" {

ga_status_t ga_params_preset(const char* preset, ga_params_t* p)
{
	if (!p) return GA_EBADARG;
	std::memset(p, 0, sizeof(*p));
	// defaults (src/AlignerMain.cpp:119-152)
	p->bandwidth = 0;
	p->maxCellsPerSlice = 0;
	p->preciseClippingIdentity = 0.66;
	p->xdrop = -1;
	p->tryAllSeeds = 0;
	p->seedExtendDensity = 0.002;
	p->seedClusterMinSize = 1;
	p->multimapScoreFraction = 0.9;
	p->minAlignmentScore = 0;
	p->minimizerLength = 19;
	p->minimizerWindowSize = 30;
	p->minimizerSeedDensity = 0;
	p->minimizerDiscardMostNumerousFraction = 0.0002;
	p->cigarMatchMismatchMerge = 0;
	p->maxSpeculativeSeeds = 64;
	p->deviceSeedOrder = 0;
	if (!preset) return GA_OK;
	std::string s = preset;
	if (s == "dbg")
	{
		p->minimizerSeedDensity = 5;
		p->minimizerLength = 19;
		p->minimizerWindowSize = 30;
		p->seedExtendDensity = 0.002;
		p->bandwidth = 5;
		p->maxCellsPerSlice = 10000;
	}
	else if (s == "vg")
	{
		p->minimizerSeedDensity = 10;
		p->minimizerLength = 15;
		p->minimizerWindowSize = 20;
		p->seedExtendDensity = -1;
		p->minimizerDiscardMostNumerousFraction = 0.001;
		p->bandwidth = 10;
	}
	else
	{
		return GA_EBADARG;
	}
	return GA_OK;
}

ga_status_t ga_graph_load_gfa(const char* path, ga_graph_t** graph, char* err, size_t errcap)
{
	if (!path || !graph) return GA_EBADARG;
	try
	{
		auto* g = new ga_graph;
		g->host = loadGfaGraph(path);
		*graph = g;
		return GA_OK;
	}
	catch (const std::bad_alloc&)
	{
		copyErr(err, errcap, "out of memory");
		return GA_ENOMEM;
	}
	catch (const std::exception& e)
	{
		copyErr(err, errcap, e.what());
		return GA_EIO;
	}
}

void ga_graph_free(ga_graph_t* graph) { delete graph; }

int64_t ga_graph_num_nodes(const ga_graph_t* g) { return g ? (int64_t)g->host.numNodes() : -1; }
int64_t ga_graph_num_edges(const ga_graph_t* g) { return g ? (int64_t)g->host.outAdj.size() : -1; }
int64_t ga_graph_first_ambiguous(const ga_graph_t* g) { return g ? (int64_t)g->host.firstAmbiguous : -1; }

ga_status_t ga_graph_export(const ga_graph_t* graph, int64_t* nodeLength, int64_t* nodeIds, int64_t* nodeOffset, int64_t* component, uint8_t* linearizable, uint8_t* reverse, int64_t* chain, int64_t* chainPos, uint64_t* chunks, int64_t* inOffsets, int64_t* inNeighbors, int64_t* outOffsets, int64_t* outNeighbors)
{
	if (!graph) return GA_EBADARG;
	const GraphHost& g = graph->host;
	size_t n = g.numNodes();
	for (size_t i = 0; i < n; i++)
	{
		nodeLength[i] = g.nodeLen[i];
		nodeIds[i] = g.nodeId[i];
		nodeOffset[i] = g.nodeOff[i];
		component[i] = g.comp[i];
		linearizable[i] = g.linearizable[i];
		reverse[i] = g.reverse[i];
		chain[i] = g.chain[i];
		chainPos[i] = (int64_t)g.chainPos[i];
		chunks[2 * i] = i < g.firstAmbiguous ? g.seq2[2 * i] : 0;
		chunks[2 * i + 1] = i < g.firstAmbiguous ? g.seq2[2 * i + 1] : 0;
	}
	for (size_t i = 0; i <= n; i++)
	{
		inOffsets[i] = g.inOff[i];
		outOffsets[i] = g.outOff[i];
	}
	for (size_t i = 0; i < g.inAdj.size(); i++) inNeighbors[i] = g.inAdj[i];
	for (size_t i = 0; i < g.outAdj.size(); i++) outNeighbors[i] = g.outAdj[i];
	return GA_OK;
}

ga_status_t ga_context_create(int device, ga_context_t** ctx, char* err, size_t errcap)
{
	if (!ctx) return GA_EBADARG;
	try
	{
		auto* c = new ga_context;
		c->ctx = std::make_unique<Context>(device);
		*ctx = c;
		return GA_OK;
	}
	catch (const HipError& e)
	{
		copyErr(err, errcap, e.what());
		int count = 0;
		if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GA_ENODEVICE;
		return GA_EHIP;
	}
	catch (const std::exception& e)
	{
		copyErr(err, errcap, e.what());
		return GA_EHIP;
	}
}

void ga_context_free(ga_context_t* ctx) { delete ctx; }

// largest bandwidth a schedule may give a slice (keeps the band's int32 score
// arithmetic, minimum + bandwidth, far from overflow)
constexpr int32_t kMaxScheduledBandwidth = 1 << 20;

static ga_status_t extendBatch(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, const int32_t* schedule, const int64_t* scheduleOffsets, const int64_t* scheduleLens, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs)
{
	if (!ctx || !graph || !params || n < 0) return GA_EBADARG;
	if (schedule && (!scheduleOffsets || !scheduleLens)) return GA_EBADARG;
	const GraphHost& g = graph->host;
	std::vector<ExtJob> jobs(n);
	for (int64_t i = 0; i < n; i++)
	{
		auto it = g.lookup.find(bigraphNodeIds[i]);
		if (it == g.lookup.end() || offsets[i] < 0 || (uint64_t)offsets[i] >= g.originalSize.at(bigraphNodeIds[i])) return GA_EBADARG;
		if (lens[i] <= 0 || lens[i] > (1 << 24) || traceCap < 2 * lens[i] + 256) return GA_EBADARG;
		uint32_t node = g.unitigNode(bigraphNodeIds[i], offsets[i]);
		jobs[i] = ExtJob{seqs[i], (uint32_t)lens[i], node, (uint32_t)(offsets[i] - g.nodeOff[node])};
		if (schedule && scheduleLens[i] > 0)
		{
			if (scheduleOffsets[i] < 0 || scheduleLens[i] > (1 << 24)) return GA_EBADARG;
			const int32_t* b = schedule + scheduleOffsets[i];
			for (int64_t k = 0; k < scheduleLens[i]; k++)
				if (b[k] < 0 || b[k] > kMaxScheduledBandwidth) return GA_EBADARG;
			jobs[i].bands = b;
			jobs[i].bandLen = (uint32_t)scheduleLens[i];
		}
	}
	try
	{
		ExtBatch out;
		ctx->ctx->extend(g, extParamsOf(*params), jobs, summaries != nullptr, out);
		ga_status_t rc = GA_OK;
		std::vector<uint64_t> items;
		for (int64_t i = 0; i < n; i++)
		{
			const auto& r = out.results[i];
			if (r.status == gad::EXT_OVERFLOW) rc = GA_EOVERFLOW;
			status[i] = r.status;
			scores[i] = r.score;
			ntrace[i] = r.ntrace;
			if (nslices) nslices[i] = r.nslices;
			int64_t* t = traces + 4 * i * traceCap;
			out.items(g, i, items);
			for (uint32_t k = 0; k < r.ntrace; k++)
			{
				uint64_t v = items[k];
				uint32_t hi = v >> 32;
				t[4 * k] = (uint32_t)v;
				t[4 * k + 1] = (hi >> 1) & 63;
				t[4 * k + 2] = (int64_t)(hi >> 7) - 1;
				t[4 * k + 3] = hi & 1;
			}
			if (summaries && r.status != gad::EXT_OVERFLOW)
				for (uint32_t k = 0; k < r.nslices && (int64_t)k < sumCap; k++)
					std::memcpy(&summaries[i * sumCap + k], &out.summaries[out.summaryOffset[i] + k], sizeof(ga_slice_summary_t));
		}
		if (kernelMs) *kernelMs = out.kernelMs;
		return rc;
	}
	catch (const HipError& e)
	{
		return GA_EHIP;
	}
	catch (const std::bad_alloc&)
	{
		return GA_ENOMEM;
	}
}

ga_status_t ga_extend_batch(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs)
{
	return extendBatch(ctx, graph, params, n, seqs, lens, bigraphNodeIds, offsets, nullptr, nullptr, nullptr, traceCap, traces, ntrace, scores, status, summaries, sumCap, nslices, kernelMs);
}

ga_status_t ga_extend_batch_scheduled(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, const int32_t* schedule, const int64_t* scheduleOffsets, const int64_t* scheduleLens, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs)
{
	if (!schedule || !scheduleOffsets || !scheduleLens) return GA_EBADARG;
	return extendBatch(ctx, graph, params, n, seqs, lens, bigraphNodeIds, offsets, schedule, scheduleOffsets, scheduleLens, traceCap, traces, ntrace, scores, status, summaries, sumCap, nslices, kernelMs);
}

ga_status_t ga_index_build(const ga_graph_t* graph, const ga_params_t* params, ga_index_t** index, char* err, size_t errcap)
{
	if (!graph || !params || !index) return GA_EBADARG;
	*index = nullptr;
	if (params->minimizerLength < 1 || params->minimizerLength > 31 || params->minimizerWindowSize < params->minimizerLength)
	{
		copyErr(err, errcap, "minimizer length must be 1..31 and not above the window size");
		return GA_EBADARG;
	}
	try
	{
		auto* x = new ga_index{buildMinimizerIndex(graph->host, params->minimizerLength, params->minimizerWindowSize, params->minimizerDiscardMostNumerousFraction), graph};
		*index = x;
		return GA_OK;
	}
	catch (const std::bad_alloc&)
	{
		copyErr(err, errcap, "out of host memory");
		return GA_ENOMEM;
	}
	catch (const std::exception& e)
	{
		copyErr(err, errcap, e.what());
		return GA_EBADARG;
	}
}

ga_status_t ga_index_build_gpu(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, ga_index_t** index, char* err, size_t errcap)
{
	if (!ctx || !graph || !params || !index) return GA_EBADARG;
	*index = nullptr;
	if (params->minimizerLength < 1 || params->minimizerLength > 31 || params->minimizerWindowSize < params->minimizerLength)
	{
		copyErr(err, errcap, "minimizer length must be 1..31 and not above the window size");
		return GA_EBADARG;
	}
	try
	{
		auto* x = new ga_index{buildMinimizerIndexGpu(*ctx->ctx, graph->host, params->minimizerLength, params->minimizerWindowSize, params->minimizerDiscardMostNumerousFraction), graph};
		*index = x;
		return GA_OK;
	}
	catch (const HipError& e)
	{
		copyErr(err, errcap, e.what());
		return GA_EHIP;
	}
	catch (const std::bad_alloc&)
	{
		copyErr(err, errcap, "out of host memory");
		return GA_ENOMEM;
	}
	catch (const std::exception& e)
	{
		copyErr(err, errcap, e.what());
		return GA_EBADARG;
	}
}

ga_status_t ga_index_export(const ga_index_t* index, uint64_t* kmers, int64_t* starts, int64_t* counts, uint64_t* positions)
{
	if (!index) return GA_EBADARG;
	const ga::MinimizerIndex& x = index->idx;
	std::vector<const ga::KmerSlot*> used;
	used.reserve(x.numKmers);
	for (const auto& slot : x.table)
		if (slot.kmer != ~0ull) used.push_back(&slot);
	std::sort(used.begin(), used.end(), [](const ga::KmerSlot* a, const ga::KmerSlot* b) { return a->kmer < b->kmer; });
	for (size_t i = 0; i < used.size(); i++)
	{
		if (kmers) kmers[i] = used[i]->kmer;
		if (starts) starts[i] = used[i]->start;
		if (counts) counts[i] = used[i]->count;
	}
	if (positions) std::copy(x.positions.begin(), x.positions.end(), positions);
	return GA_OK;
}

int64_t ga_index_num_positions(const ga_index_t* index) { return index ? (int64_t)index->idx.positions.size() : -1; }

void ga_index_free(ga_index_t* index) { delete index; }

int64_t ga_index_max_count(const ga_index_t* index) { return index ? (int64_t)index->idx.maxCount : -1; }

int64_t ga_index_num_kmers(const ga_index_t* index) { return index ? (int64_t)index->idx.numKmers : -1; }

static bool readArgsOk(int64_t n, const char* const* seqs, const int64_t* lens)
{
	if (n < 0 || (n > 0 && (!seqs || !lens))) return false;
	for (int64_t i = 0; i < n; i++)
		if (lens[i] < 0 || lens[i] > (1 << 24) || (lens[i] > 0 && !seqs[i])) return false;
	return true;
}

ga_status_t ga_seeds(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, double density, int32_t order, int64_t n, const char* const* seqs, const int64_t* lens, int64_t** seeds, int64_t* counts)
{
	if (!ctx || !graph || !index || !seeds || !counts || index->graph != graph || !readArgsOk(n, seqs, lens)) return GA_EBADARG;
	*seeds = nullptr;
	ctx->residentFromUpload = false;
	try
	{
		std::vector<const char*> s(seqs, seqs + n);
		std::vector<uint32_t> l(n);
		for (int64_t i = 0; i < n; i++) l[i] = (uint32_t)lens[i];
		std::vector<std::vector<SeedHit>> out;
		std::vector<int32_t> orderStatus;
		seedReads(*ctx->ctx, graph->host, index->idx, density, order < 0 ? 0 : order > 2 ? 2 : order, s, l, out, nullptr, &orderStatus);
		// a read whose OrderSeeds fires a reference assertion is dropped there
		for (int64_t i = 0; i < n; i++)
			if (orderStatus[i] != 0) out[i].clear();
		size_t total = 0;
		for (auto& v : out) total += v.size();
		int64_t* buf = (int64_t*)std::malloc(std::max<size_t>(1, total) * 10 * sizeof(int64_t));
		if (!buf) return GA_ENOMEM;
		int64_t* o = buf;
		for (int64_t i = 0; i < n; i++)
		{
			counts[i] = out[i].size();
			for (const SeedHit& h : out[i])
			{
				o[0] = h.nodeID;
				o[1] = h.nodeOffset;
				o[2] = h.seqPos;
				o[3] = h.matchLen;
				o[4] = h.reverse;
				o[5] = h.alignmentGraphNodeId;
				o[6] = h.alignmentGraphNodeOffset;
				o[7] = h.rawSeedGoodness;
				o[8] = h.seedGoodness;
				o[9] = h.seedClusterSize;
				o += 10;
			}
		}
		*seeds = buf;
		return GA_OK;
	}
	catch (const HipError&)
	{
		return GA_EHIP;
	}
	catch (const std::bad_alloc&)
	{
		return GA_ENOMEM;
	}
	catch (const std::exception&)
	{
		return GA_EBADARG;
	}
}

} // extern "C"

namespace {

AlignOptions alignOptionsOf(const ga_params_t& p, const ga_context* ctx = nullptr)
{
	AlignOptions opt;
	if (ctx && ctx->proposal.scorer) opt.proposal = &ctx->proposal;
	opt.ext = extParamsOf(p);
	opt.xscoreErrorCost = errorCostOf(p);
	opt.seedExtendDensity = p.seedExtendDensity;
	opt.multimapScoreFraction = p.multimapScoreFraction;
	opt.minAlignmentScore = p.minAlignmentScore;
	opt.minimizerSeedDensity = p.minimizerSeedDensity;
	opt.sloppyOptimizations = p.tryAllSeeds == 0;
	opt.minSeedClusterSize = p.seedClusterMinSize < 0 ? 0 : (uint64_t)p.seedClusterMinSize;
	opt.cigarMatchMismatchMerge = p.cigarMatchMismatchMerge != 0;
	opt.deviceSeedLoop = p.deviceSeedLoop != 0;
	opt.batchedDeviceLoop = p.deviceSeedLoop == 2;
	opt.deviceSeedOrder = p.deviceSeedOrder != 0;
	// the batched form bounds each batch's output pools (about 100 bytes per read
	// base) and replays its graph once per batch
	if (opt.batchedDeviceLoop) opt.readsPerBatch = 16384;
	opt.maxSpeculativeSeeds = p.maxSpeculativeSeeds < 1 ? 1 : (uint32_t)p.maxSpeculativeSeeds;
	if (const char* v = std::getenv("GA_PIPELINE_GROUPS"))
	{
		int x = std::atoi(v);
		if (x >= 1 && x <= 64) opt.pipelineGroups = (uint32_t)x;
	}
	if (const char* v = std::getenv("GA_GROUP_BY_LENGTH")) opt.groupByLength = std::atoi(v) != 0;
	return opt;
}

// The first n outcomes as one NUL-terminated buffer from malloc, plus statuses.
ga_status_t packOutcomes(const std::vector<ReadOutcome>& out, size_t n, const AlignStats& st, char** gaf, int32_t* status, ga_align_stats_t* stats)
{
	std::vector<size_t> at(n + 1, 0);
	for (size_t i = 0; i < n; i++)
	{
		at[i + 1] = at[i] + out[i].gaf.size();
		status[i] = out[i].status;
	}
	// large buffers on 2 MiB boundaries, marked for transparent huge pages: the
	// copy below touches every page of a few hundred MB once (freed with free())
	const size_t bytes = at.back() + 1;
	char* buf = nullptr;
	if (bytes >= ((size_t)8 << 20))
	{
		void* p = nullptr;
		if (posix_memalign(&p, (size_t)2 << 20, bytes) == 0)
		{
			buf = (char*)p;
			madvise(p, bytes & ~(((size_t)2 << 20) - 1), MADV_HUGEPAGE);
		}
	}
	else
		buf = (char*)std::malloc(bytes);
	if (!buf) return GA_ENOMEM;
	// the text runs to hundreds of MB for long reads: copy it (and take the
	// fresh pages' first-touch faults) on several threads
	const size_t parts = std::min<size_t>(16, std::max<size_t>(1, at.back() >> 24));
	std::vector<std::thread> pool;
	for (size_t t = 0; t < parts; t++)
		pool.emplace_back([&, t]() {
			for (size_t i = t; i < n; i += parts) std::memcpy(buf + at[i], out[i].gaf.data(), out[i].gaf.size());
		});
	for (auto& th : pool) th.join();
	buf[at.back()] = 0;
	*gaf = buf;
	if (stats)
	{
		stats->lookupMs = st.lookupMs;
		stats->extendKernelMs = st.extendKernelMs;
		stats->hostMs = st.hostMs;
		stats->totalMs = st.totalMs;
		stats->rounds = st.rounds;
		stats->launches = st.launches;
		stats->extensions = st.extensions;
		stats->wordSteps = st.wordSteps;
		stats->seeds = st.seeds;
		stats->reruns = st.reruns;
		stats->lookupCallMs = st.lookupCallMs;
		stats->extendCallMs = st.extendCallMs;
		stats->unusedExtensions = st.unusedExtensions;
		stats->seedingMs = st.seedingMs;
		stats->pickMs = st.pickMs;
		stats->replayMs = st.replayMs;
		stats->gafMs = st.gafMs;
		stats->gafBytes = (int64_t)at.back();
		stats->persistentWorkers = (int64_t)st.workers;
		stats->tuneMs = st.tuneMs;
		stats->proposalClusters = (int64_t)st.proposalClusters;
		stats->prunedClusters = (int64_t)st.prunedClusters;
		stats->prunedSeeds = (int64_t)st.prunedSeeds;
		stats->proposalMs = st.proposalMs;
	}
	return GA_OK;
}

template <typename F>
ga_status_t guarded(F fn)
{
	try
	{
		return fn();
	}
	catch (const HipError&)
	{
		return GA_EHIP;
	}
	catch (const std::bad_alloc&)
	{
		return GA_ENOMEM;
	}
	catch (const std::exception&)
	{
		return GA_EBADARG;
	}
}

} // namespace

extern "C" {

ga_status_t ga_align_reads(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t n, const char* const* names, const char* const* seqs, const int64_t* lens, char** gaf, int32_t* status, ga_align_stats_t* stats)
{
	if (!ctx || !graph || !index || !params || !gaf || !status || index->graph != graph || !readArgsOk(n, seqs, lens) || (n > 0 && !names)) return GA_EBADARG;
	for (int64_t i = 0; i < n; i++)
		if (!names[i]) return GA_EBADARG;
	*gaf = nullptr;
	ctx->residentFromUpload = false;
	return guarded([&]() {
		std::vector<const char*> nm(names, names + n), s(seqs, seqs + n);
		std::vector<uint32_t> l(n);
		for (int64_t i = 0; i < n; i++) l[i] = (uint32_t)lens[i];
		std::vector<ReadOutcome>& out = ctx->outcomes;
		AlignStats st;
		alignReads(*ctx->ctx, graph->host, index->idx, alignOptionsOf(*params, ctx), nm, s, l, out, st);
		return packOutcomes(out, (size_t)n, st, gaf, status, stats);
	});
}

ga_status_t ga_context_synchronize(ga_context_t* ctx)
{
	if (!ctx) return GA_EBADARG;
	return guarded([&]() {
		GA_HIP(hipSetDevice(ctx->ctx->device()));
		GA_HIP(hipDeviceSynchronize());
		return GA_OK;
	});
}

ga_status_t ga_reads_upload(ga_context_t* ctx, int64_t n, const char* const* seqs, const int64_t* lens)
{
	if (!ctx || !readArgsOk(n, seqs, lens)) return GA_EBADARG;
	ctx->residentFromUpload = false;
	return guarded([&]() {
		std::vector<const char*> s(seqs, seqs + n);
		std::vector<uint32_t> l(n);
		for (int64_t i = 0; i < n; i++) l[i] = (uint32_t)lens[i];
		ctx->ctx->uploadReads(s, l);
		ctx->residentFromUpload = true;
		return GA_OK;
	});
}

ga_status_t ga_align_resident(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t first, int64_t count, const char* const* names, char** gaf, int32_t* status, ga_align_stats_t* stats)
{
	if (!ctx || !graph || !index || !params || !gaf || !status || index->graph != graph || first < 0 || count < 0) return GA_EBADARG;
	if (!ctx->residentFromUpload) return GA_ESTALE;
	if ((uint64_t)(first + count) > ctx->ctx->reads().size() || (count > 0 && !names)) return GA_EBADARG;
	for (int64_t i = 0; i < count; i++)
		if (!names[i]) return GA_EBADARG;
	*gaf = nullptr;
	return guarded([&]() {
#ifdef GA_DEBUG_LOG
		const bool log = std::getenv("GA_ROUND_LOG") != nullptr;
#else
		constexpr bool log = false;
#endif
		auto t0 = std::chrono::steady_clock::now();
		auto since = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
		std::vector<ReadOutcome>& out = ctx->outcomes;
		if (out.size() < (size_t)count) out.resize(count);
		AlignStats st;
		alignResident(*ctx->ctx, graph->host, index->idx, alignOptionsOf(*params, ctx), names, (uint32_t)first, (uint32_t)count, out.data(), st);
		if (log) std::fprintf(stderr, "[capi] aligned at %.1f ms\n", since());
		ga_status_t rc = packOutcomes(out, (size_t)count, st, gaf, status, stats);
		if (log) std::fprintf(stderr, "[capi] packed at %.1f ms\n", since());
		return rc;
	});
}

static bool seedNodeOk(const GraphHost& g, const ga_seed_hit_t& s)
{
	if (s.nodeID < 0 || s.nodeID > INT32_MAX / 2 || s.nodeOffset < 0 || s.seqPos < 0 || s.matchLen < 0) return false;
	const int forwardId = (int)(s.reverse ? s.nodeID * 2 + 1 : s.nodeID * 2);
	auto it = g.originalSize.find(forwardId);
	if (it == g.originalSize.end() || (uint64_t)s.nodeOffset >= it->second) return false;
	if (s.alignmentGraphNodeId >= 0 && (uint64_t)s.alignmentGraphNodeId >= g.numNodes()) return false;
	if (s.alignmentGraphNodeId >= 0 && (s.alignmentGraphNodeOffset < 0 || (uint64_t)s.alignmentGraphNodeOffset >= g.nodeLen[s.alignmentGraphNodeId])) return false;
	return true;
}

static SeedHit seedHitOf(const ga_seed_hit_t& s)
{
	SeedHit h;
	h.nodeID = (int)s.nodeID;
	h.nodeOffset = (uint64_t)s.nodeOffset;
	h.seqPos = (uint64_t)s.seqPos;
	h.matchLen = (uint64_t)s.matchLen;
	h.reverse = s.reverse != 0;
	h.alignmentGraphNodeId = s.alignmentGraphNodeId < 0 ? ~0ull : (uint64_t)s.alignmentGraphNodeId;
	h.alignmentGraphNodeOffset = s.alignmentGraphNodeOffset < 0 ? ~0ull : (uint64_t)s.alignmentGraphNodeOffset;
	h.rawSeedGoodness = (uint64_t)s.rawSeedGoodness;
	h.seedGoodness = (uint64_t)s.seedGoodness;
	h.seedClusterSize = (uint64_t)s.seedClusterSize;
	return h;
}

static void toSeedHit(const SeedHit& h, ga_seed_hit_t& s)
{
	s.nodeID = h.nodeID;
	s.nodeOffset = (int64_t)h.nodeOffset;
	s.seqPos = (int64_t)h.seqPos;
	s.matchLen = (int64_t)h.matchLen;
	s.reverse = h.reverse ? 1 : 0;
	s.alignmentGraphNodeId = h.alignmentGraphNodeId == ~0ull ? -1 : (int64_t)h.alignmentGraphNodeId;
	s.alignmentGraphNodeOffset = h.alignmentGraphNodeOffset == ~0ull ? -1 : (int64_t)h.alignmentGraphNodeOffset;
	s.rawSeedGoodness = (int64_t)h.rawSeedGoodness;
	s.seedGoodness = (int64_t)h.seedGoodness;
	s.seedClusterSize = (int64_t)h.seedClusterSize;
}

ga_status_t ga_order_seeds(const ga_graph_t* graph, ga_seed_hit_t* seeds, int64_t n)
{
	if (!graph || n < 0 || (n > 0 && !seeds)) return GA_EBADARG;
	for (int64_t i = 0; i < n; i++)
		if (!seedNodeOk(graph->host, seeds[i])) return GA_EBADARG;
	return guarded([&]() {
		std::vector<SeedHit> v(n);
		for (int64_t i = 0; i < n; i++) v[i] = seedHitOf(seeds[i]);
		if (!orderSeedsByChaining(graph->host, v)) return GA_ASSERT;
		for (int64_t i = 0; i < n; i++) toSeedHit(v[i], seeds[i]);
		return GA_OK;
	});
}

ga_status_t ga_align_one_way(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const ga_seed_hit_t* seeds, const int64_t* seedCounts, ga_alignment_result_t** result)
{
	if (!ctx || !graph || !params || !result || !readArgsOk(n, seqs, lens) || (n > 0 && !seedCounts)) return GA_EBADARG;
	*result = nullptr;
	int64_t total = 0;
	for (int64_t i = 0; i < n; i++)
	{
		if (seedCounts[i] < 0) return GA_EBADARG;
		total += seedCounts[i];
	}
	if (total > 0 && !seeds) return GA_EBADARG;
	for (int64_t k = 0; k < total; k++)
		if (!seedNodeOk(graph->host, seeds[k])) return GA_EBADARG;
	ctx->residentFromUpload = false;
	return guarded([&]() {
		std::unique_ptr<ga_alignment_result> res(new ga_alignment_result);
		std::vector<std::vector<SeedHit>> perRead(n);
		int64_t at = 0;
		for (int64_t i = 0; i < n; i++)
		{
			perRead[i].resize(seedCounts[i]);
			for (int64_t k = 0; k < seedCounts[i]; k++) perRead[i][k] = seedHitOf(seeds[at + k]);
			at += seedCounts[i];
		}
		std::vector<const char*> s(seqs, seqs + n);
		std::vector<uint32_t> l(n);
		for (int64_t i = 0; i < n; i++) l[i] = (uint32_t)lens[i];
		ctx->ctx->uploadReads(s, l);
		std::vector<ReadOutcome> out(n);
		res->alignments.resize(n);
		AlignStats st;
		alignOneWayResident(*ctx->ctx, graph->host, alignOptionsOf(*params), 0, (uint32_t)n, perRead.data(), res->alignments.data(), out.data(), st);
		res->status.resize(n);
		res->seedsExtended.resize(n);
		for (int64_t i = 0; i < n; i++)
		{
			res->status[i] = out[i].status;
			res->seedsExtended[i] = out[i].seedsExtended;
		}
		*result = res.release();
		return GA_OK;
	});
}

static bool resultReadOk(const ga_alignment_result_t* r, int64_t read) { return r && read >= 0 && (uint64_t)read < r->status.size(); }

int64_t ga_result_reads(const ga_alignment_result_t* r) { return r ? (int64_t)r->status.size() : -1; }

int32_t ga_result_status(const ga_alignment_result_t* r, int64_t read) { return resultReadOk(r, read) ? r->status[read] : -1; }

int64_t ga_result_seeds_extended(const ga_alignment_result_t* r, int64_t read) { return resultReadOk(r, read) ? (int64_t)r->seedsExtended[read] : -1; }

int64_t ga_result_count(const ga_alignment_result_t* r, int64_t read) { return resultReadOk(r, read) ? (int64_t)r->alignments[read].size() : -1; }

ga_status_t ga_result_item(const ga_alignment_result_t* r, int64_t read, int64_t i, ga_alignment_item_t* item)
{
	if (!resultReadOk(r, read) || i < 0 || (uint64_t)i >= r->alignments[read].size() || !item) return GA_EBADARG;
	const Alignment& a = r->alignments[read][i];
	item->alignmentStart = (int64_t)a.alignmentStart;
	item->alignmentEnd = (int64_t)a.alignmentEnd;
	item->alignmentScore = (int64_t)a.alignmentScore;
	item->alignmentXScore = a.alignmentXScore;
	item->seedGoodness = (int64_t)a.seedGoodness;
	item->mappingQuality = a.mappingQuality;
	item->pad = 0;
	item->traceLength = (int64_t)a.trace.size();
	return GA_OK;
}

ga_status_t ga_result_trace(const ga_alignment_result_t* r, int64_t read, int64_t i, ga_trace_item_t* items, int64_t cap)
{
	if (!resultReadOk(r, read) || i < 0 || (uint64_t)i >= r->alignments[read].size()) return GA_EBADARG;
	const auto& t = r->alignments[read][i].trace;
	if (cap < (int64_t)t.size() || (!items && !t.empty())) return GA_EBADARG;
	for (size_t k = 0; k < t.size(); k++)
	{
		ga_trace_item_t& o = items[k];
		std::memset(&o, 0, sizeof(o));
		o.node = t[k].node;
		o.nodeOffset = t[k].nodeOffset;
		o.seqPos = t[k].seqPos;
		o.nodeSwitch = t[k].nodeSwitch ? 1 : 0;
		o.sequenceCharacter = t[k].seqChar;
		o.graphCharacter = t[k].graphChar;
	}
	return GA_OK;
}

ga_status_t ga_select_alignments(ga_alignment_result_t* r, int64_t read, const ga_params_t* params, int32_t sortOutput)
{
	if (!resultReadOk(r, read) || !params) return GA_EBADARG;
	return guarded([&]() {
		auto& v = r->alignments[read];
		if (!v.empty()) v = selectAlignments(v, params->minAlignmentScore, params->multimapScoreFraction);
		if (sortOutput) ga::sortForOutput(v);
		return GA_OK;
	});
}

ga_status_t ga_add_mapping_qualities(ga_alignment_result_t* r, int64_t read)
{
	if (!resultReadOk(r, read)) return GA_EBADARG;
	return addMappingQualities(r->alignments[read]) ? GA_OK : GA_ASSERT;
}

ga_status_t ga_gaf_line(const ga_graph_t* graph, const ga_alignment_result_t* r, int64_t read, int64_t i, const char* name, int64_t readLength, int32_t cigarMatchMismatchMerge, char* buf, int64_t cap, int64_t* needed)
{
	if (!graph || !resultReadOk(r, read) || i < 0 || (uint64_t)i >= r->alignments[read].size() || !name || readLength < 0 || !needed) return GA_EBADARG;
	const Alignment& a = r->alignments[read][i];
	if (a.trace.empty()) return GA_EBADARG;
	return guarded([&]() {
		std::string line;
		if (!gafLine(line, graph->host, name, (uint64_t)readLength, a, cigarMatchMismatchMerge != 0)) return GA_ASSERT;
		*needed = (int64_t)line.size() + 1;
		if (!buf || cap < *needed) return GA_EBADARG;
		std::memcpy(buf, line.c_str(), line.size() + 1);
		return GA_OK;
	});
}

void ga_result_free(ga_alignment_result_t* r) { delete r; }

ga_status_t ga_extend_batch_noalloc(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, double* kernelMs)
{
	if (!ctx || !graph || !params || n < 0 || (n > 0 && (!seqs || !lens || !bigraphNodeIds || !offsets || !traces || !ntrace || !scores || !status))) return GA_EBADARG;
	if ((uint64_t)n > Context::kNoAllocMaxRequests) return GA_ESCRATCH;
	const GraphHost& g = graph->host;
	// per-extension arrays of fixed size (n is bounded): no heap allocation; one
	// call at a time per process
	static std::mutex callMutex;
	static uint32_t lenv[Context::kNoAllocMaxRequests], nodev[Context::kNoAllocMaxRequests], offv[Context::kNoAllocMaxRequests];
	static gad::ExtResult resv[Context::kNoAllocMaxRequests];
	std::lock_guard<std::mutex> lock(callMutex);
	for (int64_t i = 0; i < n; i++)
	{
		auto it = g.originalSize.find(bigraphNodeIds[i]);
		if (it == g.originalSize.end() || offsets[i] < 0 || (uint64_t)offsets[i] >= it->second) return GA_EBADARG;
		if (lens[i] <= 0 || lens[i] > (1 << 24) || traceCap < 2 * lens[i] + 256 || !seqs[i]) return GA_EBADARG;
		const uint32_t node = g.unitigNode(bigraphNodeIds[i], offsets[i]);
		lenv[i] = (uint32_t)lens[i];
		nodev[i] = node;
		offv[i] = (uint32_t)(offsets[i] - g.nodeOff[node]);
	}
	return guarded([&]() {
		const gad::DevGraph* dg = ctx->ctx->residentGraph(g);
		if (!dg) return GA_EBADARG;
		if (n == 0)
		{
			if (kernelMs) *kernelMs = 0;
			return GA_OK;
		}
		if (!ctx->ctx->extendNoAlloc(*dg, extParamsOf(*params), (uint32_t)n, seqs, lenv, nodev, offv, traceCap, traces, resv, kernelMs)) return GA_ESCRATCH;
		ga_status_t rc = GA_OK;
		for (int64_t i = 0; i < n; i++)
		{
			const gad::ExtResult& r = resv[i];
			status[i] = r.status;
			scores[i] = r.score;
			ntrace[i] = r.status == gad::EXT_OVERFLOW ? 0 : r.ntrace;
			if (r.status == gad::EXT_OVERFLOW) rc = GA_EOVERFLOW;
		}
		return rc;
	});
}

int64_t ga_mlp_num_params(void) { return (int64_t)gad::kMlpParams; }

ga_status_t ga_mlp_create(ga_context_t* ctx, uint64_t seed, ga_mlp_t** mlp)
{
	if (!ctx || !mlp) return GA_EBADARG;
	*mlp = nullptr;
	return guarded([&]() {
		auto m = std::make_unique<ga_mlp>();
		m->model = std::make_unique<MlpModel>(ctx->ctx->device(), seed);
		*mlp = m.release();
		return GA_OK;
	});
}

void ga_mlp_free(ga_mlp_t* mlp) { delete mlp; }

ga_status_t ga_mlp_get_params(ga_mlp_t* mlp, float* params)
{
	if (!mlp || !params) return GA_EBADARG;
	return guarded([&]() {
		mlp->model->getWeights(params);
		return GA_OK;
	});
}

static bool mlpPaddingZero(const float* w)
{
	for (int i = gad::kMlpInUsed; i < gad::kMlpIn; i++)
		for (int j = 0; j < gad::kMlpHidden; j++)
			if (w[gad::kMlpW1 + (size_t)i * gad::kMlpHidden + j] != 0.0f) return false;
	for (int i = 0; i < gad::kMlpHidden; i++)
		for (int j = gad::kMlpOutUsed; j < gad::kMlpOut; j++)
			if (w[gad::kMlpW3 + (size_t)i * gad::kMlpOut + j] != 0.0f) return false;
	for (int j = gad::kMlpOutUsed; j < gad::kMlpOut; j++)
		if (w[gad::kMlpB3 + j] != 0.0f) return false;
	for (size_t k = 0; k < gad::kMlpParams; k++)
		if (!std::isfinite(w[k])) return false;
	return true;
}

ga_status_t ga_mlp_set_params(ga_mlp_t* mlp, const float* params)
{
	if (!mlp || !params || !mlpPaddingZero(params)) return GA_EBADARG;
	return guarded([&]() {
		mlp->model->setWeights(params);
		return GA_OK;
	});
}

static bool finiteRows(const float* x, int64_t n, int width)
{
	for (int64_t k = 0; k < n * width; k++)
		if (!std::isfinite(x[k])) return false;
	return true;
}

ga_status_t ga_mlp_forward(ga_mlp_t* mlp, int64_t n, const float* features, float* probs)
{
	if (!mlp || n < 0 || (n > 0 && (!features || !probs)) || n > (int64_t)1 << 30) return GA_EBADARG;
	if (!finiteRows(features, n, GA_MLP_FEATURES)) return GA_EBADARG;
	return guarded([&]() {
		mlp->model->forward(features, (size_t)n, probs);
		return GA_OK;
	});
}

static bool labelsOk(const float* y, int64_t n)
{
	for (int64_t k = 0; k < n * GA_MLP_OUTPUTS; k++)
		if (!(y[k] >= 0.0f && y[k] <= 1.0f)) return false;
	return true;
}

ga_status_t ga_mlp_gradients(ga_mlp_t* mlp, int64_t n, const float* features, const float* labels, double* loss, float* grads)
{
	if (!mlp || n < 1 || n > (int64_t)1 << 24 || !features || !labels || !grads) return GA_EBADARG;
	if (!finiteRows(features, n, GA_MLP_FEATURES) || !labelsOk(labels, n)) return GA_EBADARG;
	return guarded([&]() {
		const double l = mlp->model->gradients(features, labels, (size_t)n, grads);
		if (loss) *loss = l;
		return GA_OK;
	});
}

ga_status_t ga_mlp_train(ga_mlp_t* mlp, int64_t n, const float* features, const float* labels, const ga_mlp_train_params_t* params, float* epochLoss, double* trainMs)
{
	if (!mlp || !params || n < 1 || n > (int64_t)1 << 30 || !features || !labels) return GA_EBADARG;
	if (params->epochs < 1 || params->epochs > 64 || params->batch < 1 || params->batch > (1 << 20)) return GA_EBADARG;
	if (!(params->lr > 0) || !(params->beta1 >= 0 && params->beta1 < 1) || !(params->beta2 >= 0 && params->beta2 < 1) || !(params->eps > 0)) return GA_EBADARG;
	if (!finiteRows(features, n, GA_MLP_FEATURES) || !labelsOk(labels, n)) return GA_EBADARG;
	return guarded([&]() {
		MlpTrainOptions o;
		o.epochs = (uint32_t)params->epochs;
		o.batch = (uint32_t)params->batch;
		o.lr = params->lr;
		o.beta1 = params->beta1;
		o.beta2 = params->beta2;
		o.eps = params->eps;
		o.seed = params->seed;
		const MlpTrainStats st = mlp->model->train(features, labels, (size_t)n, o);
		if (epochLoss)
			for (size_t e = 0; e < st.epochLoss.size(); e++) epochLoss[e] = st.epochLoss[e];
		if (trainMs) *trainMs = st.ms;
		return GA_OK;
	});
}

ga_status_t ga_cluster_records(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t first, int64_t count, ga_records_t** records)
{
	if (!ctx || !graph || !index || !params || !records || index->graph != graph || first < 0 || count < 0) return GA_EBADARG;
	*records = nullptr;
	if (!ctx->residentFromUpload) return GA_ESTALE;
	if ((uint64_t)(first + count) > ctx->ctx->reads().size()) return GA_EBADARG;
	return guarded([&]() {
		auto r = std::make_unique<ga_records>();
		AlignOptions opt = alignOptionsOf(*params);
		opt.records = &r->rec;
		std::vector<std::string> nm(count, "r");
		std::vector<const char*> names(count);
		for (int64_t i = 0; i < count; i++) names[i] = nm[i].c_str();
		std::vector<ReadOutcome>& out = ctx->outcomes;
		if (out.size() < (size_t)count) out.resize(count);
		AlignStats st;
		alignResident(*ctx->ctx, graph->host, index->idx, opt, names.data(), (uint32_t)first, (uint32_t)count, out.data(), st);
		r->rec.sort();
		*records = r.release();
		return GA_OK;
	});
}

int64_t ga_records_count(const ga_records_t* r) { return r ? (int64_t)r->rec.reads.size() : -1; }
const float* ga_records_features(const ga_records_t* r) { return r ? r->rec.features.data() : nullptr; }
const float* ga_records_labels(const ga_records_t* r) { return r ? r->rec.labels.data() : nullptr; }
const int64_t* ga_records_reads(const ga_records_t* r) { return r ? r->rec.reads.data() : nullptr; }
void ga_records_free(ga_records_t* r) { delete r; }

ga_status_t ga_context_set_proposal(ga_context_t* ctx, ga_mlp_t* mlp, double threshold, int32_t keepTop, int32_t maxClusters)
{
	if (!ctx) return GA_EBADARG;
	if (!mlp)
	{
		ctx->proposal = SeedProposal{};
		return GA_OK;
	}
	if (keepTop < 1 || maxClusters < 0 || !std::isfinite(threshold) || mlp->model->device() != ctx->ctx->device()) return GA_EBADARG;
	ctx->proposal.maxClusters = (uint32_t)maxClusters;
	ctx->proposal.scorer = mlp->model.get();
	ctx->proposal.threshold = threshold;
	ctx->proposal.keepTop = (uint32_t)keepTop;
	return GA_OK;
}

void ga_free(void* p) { std::free(p); }

} // extern "C"
