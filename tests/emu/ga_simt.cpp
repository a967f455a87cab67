// Host SIMT build of the device extension code: the GPU branch of the lane
// helpers (ballots, shuffles, per-lane registers) runs on 64 std::threads that
// meet at a barrier in every wave intrinsic. Test infrastructure only: it checks
// the exact lane code and that every lane follows the same control flow (a lane
// that diverges never reaches the barrier, which the harness reports as a hang).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <vector>
#include <cstdio>

namespace simt {
struct Wave
{
	std::mutex m;
	std::condition_variable cv;
	int arrived = 0;
	uint64_t generation = 0;
	uint64_t slot[64];
	bool timedOut = false;
	void barrier()
	{
		std::unique_lock<std::mutex> lk(m);
		uint64_t gen = generation;
		if (++arrived == 64)
		{
			arrived = 0;
			generation++;
			cv.notify_all();
			return;
		}
		if (!cv.wait_for(lk, std::chrono::seconds(20), [&] { return generation != gen; }))
		{
			timedOut = true;
			throw std::runtime_error("lane divergence: barrier timed out");
		}
	}
};
thread_local int tlsLane = 0;
thread_local Wave* tlsWave = nullptr;
inline uint64_t exchange(uint64_t v, int src)
{
	Wave& w = *tlsWave;
	w.slot[tlsLane] = v;
	w.barrier();
	uint64_t r = w.slot[src];
	w.barrier();
	return r;
}
inline uint64_t ballotImpl(bool p)
{
	Wave& w = *tlsWave;
	w.slot[tlsLane] = p ? 1 : 0;
	w.barrier();
	uint64_t m = 0;
	for (int l = 0; l < 64; l++)
		if (w.slot[l]) m |= (uint64_t)1 << l;
	w.barrier();
	return m;
}
} // namespace simt

inline int __lane_id() { return simt::tlsLane; }
inline uint64_t __ballot(int p) { return simt::ballotImpl(p != 0); }
template <typename T>
inline T __shfl(T v, int src)
{
	uint64_t bits = 0;
	std::memcpy(&bits, &v, sizeof(T));
	uint64_t r = simt::exchange(bits, src & 63);
	T out;
	std::memcpy(&out, &r, sizeof(T));
	return out;
}
template <typename T>
inline T __shfl_up(T v, unsigned int d)
{
	int src = simt::tlsLane - (int)d;
	if (src < 0) src = simt::tlsLane;
	return __shfl(v, src);
}
template <typename T>
inline T __shfl_xor(T v, int m) { return __shfl(v, simt::tlsLane ^ m); }

#define GA_SIMT_EMULATION 1
#include "../../graphaligner_amd/csrc/ga_extend_core.h"
#include "../../graphaligner_amd/csrc/ga_graph.hpp"

extern "C" {

void* simt_graph_load(const char* path) { return new ga::GraphHost(ga::loadGfaGraph(path)); }
void simt_graph_free(void* g) { delete (ga::GraphHost*)g; }

// Same contract as emu_extend; returns -1 if the lanes diverged.
int simt_extend(void* gh, int32_t bandwidth, int64_t maxCells, int32_t cost, int32_t xdrop, const char* seq, int64_t len, int bigraphNodeId, int64_t offset, int64_t* traceOut, int64_t cap, int64_t* ntrace, int32_t* score, gad::SliceSummary* summaries, int64_t scap, int64_t* nslices)
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
	gad::ArenaCaps caps;
	caps.maxSlices = len / 64 + 3;
	caps.maxItems = caps.maxSlices * 48 + 1024;
	caps.maxSlots = 4096;
	caps.maxExtras = 16384;
	caps.maxPending = 4096;
	caps.pad = 0;
	// Lanes of a real wave run in lockstep, so a store every lane makes with the
	// same address and value is seen by all lanes at once; CPU threads do not run
	// in lockstep, so each lane gets its own copy of the workspace and outputs.
	const size_t bytes = gad::arenaBytes(caps) + 256;
	std::vector<std::vector<unsigned char>> arenas(64, std::vector<unsigned char>(bytes, 0xA5)); // not zeroed, like hipMalloc memory
	uint32_t node = G.unitigNode(bigraphNodeId, offset);
	gad::ExtRequest r;
	r.seqOffset = 0;
	r.traceOffset = 0;
	r.len = len;
	r.startNode = node;
	r.startOffset = offset - G.nodeOff[node];
	r.traceCap = cap;
	std::vector<std::vector<uint64_t>> traces(64, std::vector<uint64_t>(cap + 1, 0xA5A5A5A5A5A5A5A5ull));
	std::vector<std::vector<gad::SliceSummary>> sumsAll(64, std::vector<gad::SliceSummary>(caps.maxSlices));
	std::vector<gad::ExtResult> outs(64);
	simt::Wave wave;
	std::atomic<int> failed{0};
	std::vector<std::thread> threads;
	for (int l = 0; l < 64; l++)
	{
		threads.emplace_back([&, l]() {
			simt::tlsLane = l;
			simt::tlsWave = &wave;
			try
			{
				unsigned char* base = (unsigned char*)(((uintptr_t)arenas[l].data() + 255) & ~(uintptr_t)255);
				gad::runRequest(g, p, r, seq, traces[l].data(), base, caps, sumsAll[l].data(), outs[l]);
			}
			catch (const std::exception&)
			{
				failed = 1;
				std::lock_guard<std::mutex> lk(wave.m);
				wave.generation++;
				wave.cv.notify_all();
			}
		});
	}
	for (auto& t : threads) t.join();
	if (failed) return -1;
	for (int l = 1; l < 64; l++)
		if (std::memcmp(&outs[l], &outs[0], sizeof(gad::ExtResult)) != 0) return -2;
	for (int l = 1; l < 64; l++)
		if (traces[l] != traces[0]) return -3;
	// every store of the extension code is meant to be uniform: the 64 private
	// workspaces must end up identical
	{
		const size_t used = gad::arenaBytes(caps);
		auto baseOf = [&](int l) { return (unsigned char*)(((uintptr_t)arenas[l].data() + 255) & ~(uintptr_t)255); };
		for (int l = 1; l < 64; l++)
			if (std::memcmp(baseOf(l), baseOf(0), used) != 0)
			{
				size_t first = 0;
				while (baseOf(l)[first] == baseOf(0)[first]) first++;
				fprintf(stderr, "lane %d workspace differs from lane 0 at byte %zu\n", l, first);
				return -4;
			}
	}
	const gad::ExtResult& out = outs[0];
	const auto& trace = traces[0];
	const auto& sums = sumsAll[0];
	*nslices = out.nslices;
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

// The persistent work loop of the wave-cooperative kernel: `waves` waves of 64
// lane threads; lane 0 of a wave takes the next request from a shared counter,
// the index is broadcast with a shuffle, every lane leaves when it passes n.
// Requests are extensions (seqs[i], bigraph node, offset); status and score of
// each come back. Returns -1 if any wave's lanes diverged, -2 if the lanes of
// a wave disagree on a result, else 0.
int simt_persistent(void* gh, int32_t bandwidth, int64_t maxCells, int32_t cost, int32_t xdrop, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* nodes, const int64_t* offsets, int waves, int32_t* statusOut, int32_t* scoreOut)
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
	int64_t maxLen = 1;
	for (int64_t i = 0; i < n; i++) maxLen = std::max<int64_t>(maxLen, lens[i]);
	gad::ArenaCaps caps;
	caps.maxSlices = maxLen / 64 + 3;
	caps.maxItems = caps.maxSlices * 48 + 1024;
	caps.maxSlots = 4096;
	caps.maxExtras = 16384;
	caps.maxPending = 4096;
	caps.pad = 0;
	const size_t bytes = gad::arenaBytes(caps) + 256;
	const int64_t cap = 4 * maxLen + 1024;
	std::atomic<uint32_t> counter{0};
	std::vector<simt::Wave> waveState(waves);
	std::vector<std::vector<gad::ExtResult>> perLane(waves * 64, std::vector<gad::ExtResult>(n));
	std::atomic<int> failed{0};
	std::vector<std::thread> threads;
	for (int w = 0; w < waves; w++)
		for (int l = 0; l < 64; l++)
			threads.emplace_back([&, w, l]() {
				simt::tlsLane = l;
				simt::tlsWave = &waveState[w];
				// each lane its own workspace and trace buffer (lanes are not in lockstep here)
				std::vector<unsigned char> arena(bytes, 0xA5);
				std::vector<uint64_t> trace(cap + 1, 0);
				unsigned char* base = (unsigned char*)(((uintptr_t)arena.data() + 255) & ~(uintptr_t)255);
				try
				{
					while (true)
					{
						uint32_t idx = 0;
						if (l == 0) idx = counter.fetch_add(1u);
						idx = __shfl(idx, 0);
						if (idx >= (uint64_t)n) break;
						gad::ExtRequest r;
						const uint32_t node = G.unitigNode(nodes[idx], offsets[idx]);
						r.seqOffset = 0;
						r.traceOffset = 0;
						r.len = lens[idx];
						r.startNode = node;
						r.startOffset = offsets[idx] - G.nodeOff[node];
						r.traceCap = cap;
						gad::ExtResult out;
						gad::runRequest(g, p, r, seqs[idx], trace.data(), base, caps, nullptr, out);
						perLane[w * 64 + l][idx] = out;
					}
				}
				catch (const std::exception&)
				{
					failed = 1;
					for (auto& ws : waveState)
					{
						std::lock_guard<std::mutex> lk(ws.m);
						ws.generation++;
						ws.cv.notify_all();
					}
				}
			});
	for (auto& t : threads) t.join();
	if (failed) return -1;
	// the wave that ran request i: all its lanes hold the same result
	for (int64_t i = 0; i < n; i++)
	{
		int owner = -1;
		for (int w = 0; w < waves && owner < 0; w++)
			if (perLane[w * 64][i].nslices || perLane[w * 64][i].ntrace || perLane[w * 64][i].status) owner = w;
		if (owner < 0) owner = 0;
		for (int l = 1; l < 64; l++)
			if (std::memcmp(&perLane[owner * 64 + l][i], &perLane[owner * 64][i], sizeof(gad::ExtResult)) != 0) return -2;
		statusOut[i] = perLane[owner * 64][i].status;
		scoreOut[i] = perLane[owner * 64][i].score;
	}
	return 0;
}
}
