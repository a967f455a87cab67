// extern "C" boundary (include/graphaligner_amd.h).
#include "../../include/graphaligner_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
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

ga_status_t ga_extend_batch(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs)
{
	if (!ctx || !graph || !params || n < 0) return GA_EBADARG;
	const GraphHost& g = graph->host;
	std::vector<ExtJob> jobs(n);
	for (int64_t i = 0; i < n; i++)
	{
		auto it = g.lookup.find(bigraphNodeIds[i]);
		if (it == g.lookup.end() || offsets[i] < 0 || (uint64_t)offsets[i] >= g.originalSize.at(bigraphNodeIds[i])) return GA_EBADARG;
		if (lens[i] <= 0 || lens[i] > (1 << 24) || traceCap < 2 * lens[i] + 256) return GA_EBADARG;
		uint32_t node = g.unitigNode(bigraphNodeIds[i], offsets[i]);
		jobs[i] = ExtJob{seqs[i], (uint32_t)lens[i], node, (uint32_t)(offsets[i] - g.nodeOff[node])};
	}
	try
	{
		ExtBatch out;
		ctx->ctx->extend(g, extParamsOf(*params), jobs, summaries != nullptr, out);
		ga_status_t rc = GA_OK;
		for (int64_t i = 0; i < n; i++)
		{
			const auto& r = out.results[i];
			if (r.status == gad::EXT_OVERFLOW) rc = GA_EOVERFLOW;
			status[i] = r.status;
			scores[i] = r.score;
			ntrace[i] = r.ntrace;
			if (nslices) nslices[i] = r.nslices;
			int64_t* t = traces + 4 * i * traceCap;
			for (uint32_t k = 0; k < r.ntrace; k++)
			{
				uint64_t v = out.trace[out.traceOffset[i] + k];
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

} // extern "C"
