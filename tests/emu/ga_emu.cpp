// Host build of the device extension logic (graphaligner_amd/csrc/ga_extend_core.h
// with GA_WAVE_EMULATION: lane helpers run as loops over 64 lanes). Test
// infrastructure only: lets the CPU suite compare the device algorithm with the
// oracle before it runs on a GPU. The product library never contains this build.
#ifndef GA_THREAD_PER_EXTENSION
#define GA_WAVE_EMULATION 1
#endif
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>
#include "../../graphaligner_amd/csrc/ga_extend_core.h"
#include "../../graphaligner_amd/csrc/ga_graph.hpp"

extern "C" {

void* emu_graph_load(const char* path, char* err, int cap)
{
	try
	{
		return new ga::GraphHost(ga::loadGfaGraph(path));
	}
	catch (const std::exception& e)
	{
		std::strncpy(err, e.what(), cap - 1);
		err[cap - 1] = 0;
	}
	return nullptr;
}

void emu_graph_free(void* g) { delete (ga::GraphHost*)g; }

// rows: node, nodeOffset, seqPos, nodeSwitch (same layout as the oracle's orc_extend)
int emu_extend(void* gh, int32_t bandwidth, int64_t maxCells, int32_t cost, int32_t xdrop, const char* seq, int64_t len, int bigraphNodeId, int64_t offset, int64_t* traceOut, int64_t cap, int64_t* ntrace, int32_t* score, gad::SliceSummary* summaries, int64_t scap, int64_t* nslices, uint64_t* wordSteps, const int32_t* schedule, int64_t scheduleLen)
{
	auto& G = *(ga::GraphHost*)gh;
	gad::DevGraph g;
	g.nodeLen = G.nodeLen.data();
	g.seq2 = G.seq2.data();
	g.amb4 = G.amb4.data();
	g.comp = G.comp.data();
	g.linearizable = G.linearizable.data();
	g.inOff = G.inOff.data();
	g.inAdj = G.inAdj.data();
	g.outOff = G.outOff.data();
	g.outAdj = G.outAdj.data();
	g.rec = G.nodeRec.data();
	g.outEdges = G.outEdges.data();
	g.numNodes = G.numNodes();
	g.firstAmbiguous = G.firstAmbiguous;
	gad::ExtParams p;
	p.bandwidth = bandwidth;
	p.errorCost = cost;
	p.xdrop = xdrop;
	p.pad = 0;
	p.maxCellsPerSlice = maxCells <= 0 ? INT64_MAX : maxCells;
	p.bandSched = schedule;
	gad::ArenaCaps caps;
	caps.maxSlices = len / 64 + 3;
	caps.maxItems = 200000;
	caps.maxSlots = 20000;
	caps.maxExtras = 80000;
	caps.maxPending = 20000;
	caps.pad = 0;
	std::vector<unsigned char> arena(gad::arenaBytes(caps) + 256);
	unsigned char* base = (unsigned char*)(((uintptr_t)arena.data() + 255) & ~(uintptr_t)255);
	uint32_t node = G.unitigNode(bigraphNodeId, offset);
	gad::ExtRequest r;
	r.seqOffset = 0;
	r.traceOffset = 0;
	r.len = len;
	r.startNode = node;
	r.startOffset = offset - G.nodeOff[node];
	r.traceCap = cap;
	r.bandOff = 0;
	r.bandLen = schedule ? (uint32_t)scheduleLen : 0;
	std::vector<uint64_t> trace(cap + 1);
	std::vector<gad::SliceSummary> sums(caps.maxSlices);
	gad::ExtResult out;
	gad::runRequest(g, p, r, seq, trace.data(), base, caps, sums.data(), out);
	*nslices = out.nslices;
	*wordSteps = out.wordSteps;
	if (summaries)
		for (uint32_t i = 0; i < out.nslices && (int64_t)i < scap; i++) summaries[i] = sums[i];
	*ntrace = out.ntrace;
	*score = out.score;
	for (uint32_t i = 0; i < out.ntrace; i++)
	{
		uint64_t v = trace[i];
		uint32_t hi = v >> 32;
		traceOut[4 * i] = (uint32_t)v;
		traceOut[4 * i + 1] = (hi >> 1) & 63;
		traceOut[4 * i + 2] = (int64_t)(hi >> 7) - 1;
		traceOut[4 * i + 3] = hi & 1;
	}
	return out.status;
}

// Column primitives of the device code, for property tests against brute force.
void emu_merge(uint64_t avp, uint64_t avn, int32_t aend, uint64_t bvp, uint64_t bvn, int32_t bend, uint64_t* vp, uint64_t* vn, int32_t* end)
{
	gad::Slice r = gad::lpMerge(gad::mkslice(avp, avn, aend), gad::mkslice(bvp, bvn, bend));
	*vp = r.VP;
	*vn = r.VN;
	*end = r.end;
}

uint64_t emu_less_mask(uint64_t avp, uint64_t avn, int32_t aend, uint64_t bvp, uint64_t bvn, int32_t bend)
{
	return gad::lpLessMask(gad::mkslice(avp, avn, aend), gad::mkslice(bvp, bvn, bend));
}

int32_t emu_max_x(uint64_t vp, uint64_t vn, int32_t end, int32_t seqOffset, int32_t cost, int32_t cells, uint64_t extraMask)
{
	gad::Slice s = gad::mkslice(vp, vn, end);
	return gad::lpMaxX(s, seqOffset, cost, gad::maxXCandidates(s, cells, extraMask));
}
}
