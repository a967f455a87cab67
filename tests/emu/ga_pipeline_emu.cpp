// Host build of the whole alignment pipeline for the CPU test suite: the
// product's orchestrator (ga_align.cpp), index (ga_index.cpp) and graph builder
// run unchanged over a backend that executes the device code of the gfx950
// kernels (the one-extension-per-thread extension core and the seed lookup core)
// on host threads. Test infrastructure only; the product library never contains it.
#define GA_THREAD_PER_EXTENSION 1
#define GA_THREAD_HOST_TEST 1
#include <atomic>
#include <cmath>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <random>
#include <vector>
#include "../../graphaligner_amd/csrc/ga_align.hpp"
#include "../../graphaligner_amd/csrc/ga_proposal.hpp"
#include "../../graphaligner_amd/csrc/ga_extend_core.h"
#include "../../graphaligner_amd/csrc/ga_read_core.h"
#include "../../graphaligner_amd/csrc/ga_seed_core.h"
#include "../../graphaligner_amd/csrc/ga_order_core.h"

namespace {

// failure injection for the error-path tests: 1 drops every escaped trace item
// of an extension batch (the replay then meets a trace it cannot decode), 2
// reports a seed match outside the index's position array
int g_fault = 0;

// a logistic scorer over the cluster features (both outputs the same), standing
// in for the network in CPU tests of the proposal's pruning
struct LinearScorer : ga::ClusterScorer
{
	std::vector<float> w;
	float bias = 0;
	void score(const float* f, size_t n, float* out) override
	{
		for (size_t i = 0; i < n; i++)
		{
			double z = bias;
			for (int k = 0; k < ga::kClusterFeatures; k++) z += (double)w[k] * f[i * ga::kClusterFeatures + k];
			out[i * 2] = out[i * 2 + 1] = (float)(1.0 / (1.0 + std::exp(-z)));
		}
	}
};
LinearScorer g_scorer;
ga::SeedProposal g_proposal;
bool g_proposalOn = false;
std::unique_ptr<ga::ClusterRecords> g_records;
ga::AlignStats g_lastStats;

template <typename F>
void forEachParallel(size_t n, F fn)
{
	std::atomic<size_t> next{0};
	unsigned t = getenv("PIPE_THREADS") ? (unsigned)atoi(getenv("PIPE_THREADS")) : std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
	std::vector<std::thread> pool;
	for (unsigned k = 0; k < t; k++)
		pool.emplace_back([&]() {
			while (true)
			{
				size_t i = next.fetch_add(1);
				if (i >= n) break;
				fn(i);
			}
		});
	for (auto& th : pool) th.join();
}

class HostBackend : public ga::ExtendBackend
{
public:
	gad::DevGraph view(const ga::GraphHost& G)
	{
		std::lock_guard<std::mutex> lock(viewMutex_);
		if (bigSizeUid_ != G.uid)
		{
			bigSize_.resize(G.numNodes());
			for (size_t i = 0; i < G.numNodes(); i++) bigSize_[i] = (uint32_t)G.originalSize.at(G.nodeId[i]);
			bigSizeUid_ = G.uid;
		}
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
		g.nodeId = G.nodeId.data();
		g.nodeOff = G.nodeOff.data();
		g.bigSize = bigSize_.data();
		g.numNodes = G.numNodes();
		g.firstAmbiguous = G.firstAmbiguous;
		return g;
	}

	void runReadLoops(const ga::GraphHost& G, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ga::ReadLoopJob>& jobs, ga::ReadLoopOutput& out) override
	{
		const gad::DevGraph g = view(G);
		const size_t n = jobs.size();
		out = ga::ReadLoopOutput{};
		out.results.assign(n, gad::ReadResult{});
		out.alnOffset.assign(n, 0);
		out.poolOffset.assign(n, 0);
		std::vector<std::vector<gad::DevAlignment>> alns(n);
		std::vector<std::vector<uint64_t>> pools(n);
		forEachParallel(n, [&](size_t i) {
			const ga::ReadLoopJob& j = jobs[i];
			const uint32_t len = rs_.length[j.read];
			ga::ReadLoopCaps caps = ga::initialReadLoopCaps();
			for (int attempt = 0; attempt < ga::kArenaAttempts; attempt++)
			{
				gad::ReadJob r;
				r.fwdOffset = rs_.offset[j.read];
				r.revOffset = revBase_ + rs_.offset[j.read];
				r.seedStart = 0;
				r.arenaOffset = r.traceOffset = r.poolOffset = r.alnOffset = 0;
				r.extendSeeds = j.extendSeeds;
				r.len = len;
				r.seedCount = j.seedCount;
				r.traceCap = ga::traceCapFor(len);
				r.poolCap = ga::poolCapFor(len, caps);
				r.alnCap = caps.alnCap;
				r.pad = 0;
				gad::ArenaCaps c = caps.arena;
				c.maxSlices = gad::requestSlices(len);
				c.maxItems = gad::requestItems(len, caps.arena);
				std::vector<unsigned char> arena(gad::arenaBytes(c) + 256);
				unsigned char* base = (unsigned char*)(((uintptr_t)arena.data() + 255) & ~(uintptr_t)255);
				std::vector<uint64_t> raw(2ull * r.traceCap);
				pools[i].assign(r.poolCap, 0);
				alns[i].assign(r.alnCap, gad::DevAlignment{});
				gad::ReadResult res;
				gad::runReadLoop(g, p, r, reads_.data(), seeds.data() + j.seedStart, raw.data(), pools[i].data(), alns[i].data(), base, c, res, nullptr, planes_.data());
				out.results[i] = res;
				if (res.status != gad::READLOOP_OVERFLOW) break;
				ga::growReadLoopCaps(caps);
			}
			const gad::ReadResult& res = out.results[i];
			alns[i].resize(res.status == gad::READLOOP_DONE ? res.nAlignments : 0);
			pools[i].resize(res.status == gad::READLOOP_DONE ? res.poolUsed : 0);
		});
		for (size_t i = 0; i < n; i++)
		{
			out.alnOffset[i] = out.alns.size();
			out.alns.insert(out.alns.end(), alns[i].begin(), alns[i].end());
			out.poolOffset[i] = out.pool.size();
			out.pool.insert(out.pool.end(), pools[i].begin(), pools[i].end());
			out.wordSteps += out.results[i].wordSteps;
			out.extensions += out.results[i].extensions;
		}
		out.launches++;
	}

	// the batched persistent form on host threads: three workers pull reads
	// (longest first) from a shared counter, each reusing one workspace, one
	// raw-trace region and one band-queue store for every read it takes
	void runReadLoopsBatched(const ga::GraphHost& G, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ga::ReadLoopJob>& jobs, ga::ReadLoopOutput& out) override
	{
		const gad::DevGraph g = view(G);
		const size_t n = jobs.size();
		out = ga::ReadLoopOutput{};
		out.results.assign(n, gad::ReadResult{});
		out.alnOffset.assign(n, 0);
		out.poolOffset.assign(n, 0);
		if (n == 0) return;
		std::vector<uint32_t> order(n);
		for (uint32_t i = 0; i < n; i++) order[i] = i;
		std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rs_.length[jobs[a].read] > rs_.length[jobs[b].read]; });
		ga::ReadLoopCaps caps = ga::initialReadLoopCaps();
		caps.arena = ga::initialArenaCaps();
		uint32_t maxLen = 0;
		for (const auto& j : jobs) maxLen = std::max(maxLen, rs_.length[j.read]);
		gad::ArenaCaps big = caps.arena;
		big.maxSlices = gad::requestSlices(maxLen);
		big.maxItems = gad::requestItems(maxLen, caps.arena);
		const uint64_t arenaBytes = gad::arenaBytes(big) + 256;
		std::vector<std::vector<gad::DevAlignment>> alns(n);
		std::vector<std::vector<uint64_t>> pools(n);
		std::atomic<size_t> next{0};
		const int workers = 3;
		std::vector<std::thread> pool;
		for (int w = 0; w < workers; w++)
			pool.emplace_back([&]() {
				std::vector<unsigned char> arena(arenaBytes);
				unsigned char* base = (unsigned char*)(((uintptr_t)arena.data() + 255) & ~(uintptr_t)255);
				std::vector<uint64_t> raw(2ull * ga::traceCapFor(maxLen));
				std::vector<uint64_t> queue(gad::kQueueWords);
				while (true)
				{
					const size_t k = next.fetch_add(1);
					if (k >= n) break;
					const uint32_t i = order[k];
					const ga::ReadLoopJob& j = jobs[i];
					const uint32_t len = rs_.length[j.read];
					gad::ReadJob r;
					r.fwdOffset = rs_.offset[j.read];
					r.revOffset = revBase_ + rs_.offset[j.read];
					r.seedStart = 0;
					r.arenaOffset = r.traceOffset = r.poolOffset = r.alnOffset = 0;
					r.extendSeeds = j.extendSeeds;
					r.len = len;
					r.seedCount = j.seedCount;
					r.traceCap = ga::traceCapFor(len);
					r.poolCap = ga::poolCapFor(len, caps);
					r.alnCap = caps.alnCap;
					r.pad = 0;
					gad::ArenaCaps c = caps.arena;
					c.maxSlices = gad::requestSlices(len);
					c.maxItems = gad::requestItems(len, caps.arena);
					pools[i].assign(r.poolCap, 0);
					alns[i].assign(r.alnCap, gad::DevAlignment{});
					gad::ReadResult res;
					gad::runReadLoop(g, p, r, reads_.data(), seeds.data() + j.seedStart, raw.data(), pools[i].data(), alns[i].data(), base, c, res, queue.data(), planes_.data(), false, queue.data(), true);
					out.results[i] = res;
				}
			});
		for (auto& t : pool) t.join();
		std::vector<ga::ReadLoopJob> again;
		std::vector<uint32_t> againIds;
		for (size_t i = 0; i < n; i++)
		{
			const gad::ReadResult& res = out.results[i];
			out.wordSteps += res.wordSteps;
			out.extensions += res.extensions;
			if (res.status == gad::READLOOP_OVERFLOW)
			{
				again.push_back(jobs[i]);
				againIds.push_back((uint32_t)i);
				continue;
			}
			alns[i].resize(res.status == gad::READLOOP_DONE ? res.nAlignments : 0);
			pools[i].resize(res.status == gad::READLOOP_DONE ? res.poolUsed : 0);
			out.alnOffset[i] = out.alns.size();
			out.alns.insert(out.alns.end(), alns[i].begin(), alns[i].end());
			out.poolOffset[i] = out.pool.size();
			out.pool.insert(out.pool.end(), pools[i].begin(), pools[i].end());
		}
		out.launches++;
		out.workers = workers;
		if (!again.empty())
		{
			ga::ReadLoopOutput ro;
			runReadLoops(G, p, seeds, again, ro);
			out.reruns += again.size() + ro.reruns;
			out.wordSteps += ro.wordSteps;
			out.extensions += ro.extensions;
			for (size_t k = 0; k < again.size(); k++)
			{
				const uint32_t i = againIds[k];
				out.results[i] = ro.results[k];
				out.alnOffset[i] = out.alns.size() + ro.alnOffset[k];
				out.poolOffset[i] = out.pool.size() + ro.poolOffset[k];
			}
			out.alns.insert(out.alns.end(), ro.alns.begin(), ro.alns.end());
			out.pool.insert(out.pool.end(), ro.pool.begin(), ro.pool.end());
		}
	}

	uint32_t extendLanes() const override { return 2; }

	void extend(const ga::GraphHost& G, const gad::ExtParams& p, const std::vector<ga::ExtJob>& jobs, bool, ga::ExtBatch& out, const char* devSeqBase, uint32_t lane) override
	{
		const gad::DevGraph g = view(G);
		const size_t n = jobs.size();
		out.results.assign(n, gad::ExtResult{});
		out.traceOffset.assign(n, 0);
		out.summaryOffset.assign(n, 0);
		out.summaries.clear();
		std::vector<std::vector<uint64_t>> traces(n);
		std::atomic<uint64_t> reruns{0};
		if (devSeqBase && devSeqBase != reads_.data()) throw std::runtime_error("device sequences are not the resident reads");
		forEachParallel(n, [&](size_t i) {
			const ga::ExtJob& j = jobs[i];
			gad::ArenaCaps caps = ga::initialArenaCaps();
			// planes of the resident bytes, or of the job's own sequence
			std::vector<uint64_t> ownPlanes;
			if (!devSeqBase)
			{
				ownPlanes.assign(gad::planeBlockCount(j.len) * gad::kPlaneWords, 0);
				for (uint64_t b = 0; b < gad::planeBlockCount(j.len); b++) gad::planeBlock(j.seq, j.len, b, ownPlanes.data() + b * gad::kPlaneWords);
			}
			const uint64_t* planes = devSeqBase ? planes_.data() : ownPlanes.data();
			for (int attempt = 0; attempt < ga::kArenaAttempts; attempt++)
			{
				gad::ArenaCaps c = caps;
				c.maxSlices = gad::requestSlices(j.len);
				c.maxItems = gad::requestItems(j.len, caps);
				std::vector<unsigned char> arena(gad::arenaBytes(c) + 256);
				unsigned char* base = (unsigned char*)(((uintptr_t)arena.data() + 255) & ~(uintptr_t)255);
				gad::ExtRequest r;
				r.seqOffset = devSeqBase ? j.devOffset : 0;
				r.traceOffset = 0;
				r.len = j.len;
				r.startNode = j.startNode;
				r.startOffset = j.startOffset;
				r.traceCap = ga::traceCapFor(j.len);
				traces[i].assign(r.traceCap, 0);
				gad::ExtResult res;
				gad::runRequest(g, p, r, devSeqBase ? devSeqBase : j.seq, traces[i].data(), base, c, nullptr, res, nullptr, planes);
				out.results[i] = res;
				if (res.status != gad::EXT_OVERFLOW) break;
				if (attempt + 1 < ga::kArenaAttempts) reruns++;
				ga::growArenaCaps(caps);
			}
			traces[i].resize(out.results[i].status == gad::EXT_OVERFLOW ? 0 : out.results[i].ntrace);
		});
		for (size_t i = 0; i < n; i++) out.wordSteps += out.results[i].wordSteps;
		ga::encodeTraces(G, traces, codeStore_[lane], out);
		if (g_fault == 1) out.escapes.clear();
		out.launches++;
		out.reruns = reruns.load();
	}

	void uploadReads(const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens) override
	{
		const size_t n = seqs.size();
		rs_.offset.resize(n);
		rs_.length.assign(lens.begin(), lens.end());
		rs_.bytes = 0;
		rs_.maxLen = 0;
		for (size_t i = 0; i < n; i++)
		{
			rs_.offset[i] = rs_.bytes;
			rs_.bytes += lens[i];
			rs_.maxLen = std::max(rs_.maxLen, lens[i]);
		}
		rs_.host.assign(rs_.bytes + 1, 0);
		for (size_t i = 0; i < n; i++) std::memcpy(rs_.host.data() + rs_.offset[i], seqs[i], lens[i]);
		revBase_ = rs_.bytes + 1;
		reads_.assign(2 * (rs_.bytes + 1), 0);
		std::memcpy(reads_.data(), rs_.host.data(), rs_.bytes + 1);
		computePlanes();
	}

	// match-mask planes of the resident bytes, as the device keeps them
	void computePlanes()
	{
		const uint64_t blocks = gad::planeBlockCount(reads_.size());
		planes_.assign(blocks * gad::kPlaneWords, 0);
		forEachParallel(blocks, [&](size_t b) { gad::planeBlock(reads_.data(), reads_.size(), b, planes_.data() + b * gad::kPlaneWords); });
	}

	const ga::ReadSet& reads() const override { return rs_; }

	void lookupSeeds(const ga::MinimizerIndex& idx, uint32_t first, uint32_t count, ga::SeedMatches& out) override
	{
		std::vector<gad::SeedMatch> all(rs_.bytes + 1);
		out.matchCount.assign(count, 0);
		const gad::SeedTableSlot* table = reinterpret_cast<const gad::SeedTableSlot*>(idx.table.data());
		forEachParallel(count, [&](size_t i) {
			uint64_t off = rs_.offset[first + i];
			uint32_t len = rs_.length[first + i];
			for (uint32_t k = 0; k < len; k++) reads_[revBase_ + off + k] = gad::complementByte(reads_[off + len - 1 - k]);
			out.matchCount[i] = gad::lookupReadKmers(reads_.data() + off, len, table, idx.mask, idx.k, idx.w, idx.maxCount, all.data() + off);
		});
		out.matchOffset.resize(count);
		uint64_t total = 0;
		for (size_t i = 0; i < count; i++)
		{
			out.matchOffset[i] = total;
			total += out.matchCount[i];
		}
		out.matches.resize(total);
		for (size_t i = 0; i < count; i++) std::memcpy(out.matches.data() + out.matchOffset[i], all.data() + rs_.offset[first + i], out.matchCount[i] * sizeof(gad::SeedMatch));
		if (g_fault == 2 && total > 0) out.matches[total - 1].start = (uint32_t)idx.positions.size();
		out.lookupMs = 0;
		computePlanes();
	}

	// the device seeding code (ga_order_core.h) built for the host, one read per task
	void orderedSeeds(const ga::GraphHost& G, const ga::MinimizerIndex& idx, uint32_t first, uint32_t count, const ga::SeedMatches& sm, double density,
	                  std::vector<std::vector<ga::SeedHit>>& seeds, std::vector<uint8_t>& asserted) override
	{
		seeds.assign(count, {});
		asserted.assign(count, 0);
		gad::SeedGraphView gv{G.nodeId.data(), G.nodeOff.data(), G.reverse.data(), G.chain.data(), G.chainPos.data()};
		std::atomic<bool> bad{false};
		forEachParallel(count, [&](size_t i) {
			std::vector<gad::SeedMatch> m(sm.data() + sm.matchOffset[i], sm.data() + sm.matchOffset[i] + sm.matchCount[i]);
			uint64_t cap = 0;
			for (const auto& x : m) cap += x.count;
			std::vector<gad::DevSeedHit> out(cap + 1);
			std::vector<gad::ChainKey> scratch(cap + 1);
			gad::SortRange stack[gad::kSortStackEntries];
			const int64_t n = gad::seedsOfRead(gv, idx.positions.data(), idx.positions.size(), idx.maxCount, idx.k, m.data(), (uint32_t)m.size(),
			                                   rs_.length[first + i], density, out.data(), cap, scratch.data(), stack);
			if (n == -1)
			{
				asserted[i] = 1;
				return;
			}
			if (n < 0)
			{
				bad = true;
				return;
			}
			auto& v = seeds[i];
			v.resize(n);
			for (int64_t k = 0; k < n; k++)
			{
				const gad::DevSeedHit& d = out[k];
				v[k].nodeID = d.nodeID;
				v[k].nodeOffset = d.nodeOffset;
				v[k].seqPos = d.seqPos;
				v[k].matchLen = d.matchLen;
				v[k].reverse = d.reverse != 0;
				v[k].alignmentGraphNodeId = d.alignmentGraphNodeId;
				v[k].alignmentGraphNodeOffset = d.alignmentGraphNodeOffset;
				v[k].rawSeedGoodness = d.rawSeedGoodness;
				v[k].seedGoodness = d.seedGoodness;
				v[k].seedClusterSize = d.seedClusterSize;
			}
		});
		if (bad) throw std::runtime_error("seed ordering returned an inconsistent seed count");
	}

	void prepareReads(uint32_t first, uint32_t count) override
	{
		forEachParallel(count, [&](size_t i) {
			uint64_t off = rs_.offset[first + i];
			uint32_t len = rs_.length[first + i];
			for (uint32_t k = 0; k < len; k++) reads_[revBase_ + off + k] = gad::complementByte(reads_[off + len - 1 - k]);
		});
		computePlanes();
	}

	const char* readsBase() const override { return reads_.data(); }
	uint64_t reverseBase() const override { return revBase_; }

private:
	std::vector<uint8_t> codeStore_[2];
	std::mutex viewMutex_;
	std::vector<uint32_t> bigSize_;
	uint64_t bigSizeUid_ = 0;
	ga::ReadSet rs_;
	std::vector<char> reads_;
	std::vector<uint64_t> planes_;
	uint64_t revBase_ = 0;
};

struct EmuIndex
{
	ga::MinimizerIndex idx;
	const ga::GraphHost* g;
};

} // namespace

extern "C" {

void* pipe_graph_load(const char* path, char* err, int cap)
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

void pipe_graph_free(void* g) { delete (ga::GraphHost*)g; }

void* pipe_index_build(void* g, int k, int w, double discard)
{
	return new EmuIndex{ga::buildMinimizerIndex(*(ga::GraphHost*)g, k, w, discard), (ga::GraphHost*)g};
}

void pipe_index_free(void* i) { delete (EmuIndex*)i; }

int64_t pipe_index_maxcount(void* i) { return ((EmuIndex*)i)->idx.maxCount; }

int64_t pipe_index_num_kmers(void* i) { return ((EmuIndex*)i)->idx.numKmers; }

// Seeds of one read in the layout of the oracle's orc_get_seeds.
int pipe_get_seeds(void* ih, const char* seq, int64_t len, double density, int order, int64_t* out, int64_t cap, int64_t* n)
{
	auto* x = (EmuIndex*)ih;
	HostBackend be;
	std::vector<std::vector<ga::SeedHit>> seeds;
	std::vector<int32_t> status;
	ga::seedReads(be, *x->g, x->idx, density, order, {seq}, {(uint32_t)len}, seeds, nullptr, &status);
	*n = seeds[0].size();
	if (status[0] != 0)
	{
		*n = 0;
		return 2;
	}
	if ((int64_t)seeds[0].size() > cap) return 1;
	for (size_t i = 0; i < seeds[0].size(); i++)
	{
		const auto& h = seeds[0][i];
		int64_t* o = out + 10 * i;
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
	}
	return 0;
}

// Seeds of one read through the device seeding code (ga_order_core.h) built
// for the host: lookupReadKmers, then seedsOfRead. Same layout and return codes
// as pipe_get_seeds (ordered seeds only).
int pipe_get_seeds_device_code(void* ih, const char* seq, int64_t len, double density, int64_t* out, int64_t cap, int64_t* n)
{
	auto* x = (EmuIndex*)ih;
	const ga::GraphHost& g = *x->g;
	std::vector<gad::SeedTableSlot> table(x->idx.table.size());
	for (size_t i = 0; i < table.size(); i++) table[i] = gad::SeedTableSlot{x->idx.table[i].kmer, x->idx.table[i].start, x->idx.table[i].count};
	std::vector<gad::SeedMatch> m(len + 1);
	const uint32_t nm = gad::lookupReadKmers(seq, len, table.data(), x->idx.mask, x->idx.k, x->idx.w, x->idx.maxCount, m.data());
	uint64_t bound = 0;
	for (uint32_t i = 0; i < nm; i++) bound += m[i].count;
	std::vector<gad::DevSeedHit> seeds(bound + 1);
	std::vector<gad::ChainKey> scratch(bound + 1);
	gad::SeedGraphView gv{g.nodeId.data(), g.nodeOff.data(), g.reverse.data(), g.chain.data(), g.chainPos.data()};
	gad::SortRange stack[gad::kSortStackEntries];
	const int64_t r = gad::seedsOfRead(gv, x->idx.positions.data(), x->idx.positions.size(), x->idx.maxCount, x->idx.k, m.data(), nm, len, density, seeds.data(), bound,
	                                   scratch.data(), stack);
	*n = 0;
	if (r == -1) return 2;
	if (r < 0) return 3;
	*n = r;
	if (r > cap) return 1;
	for (int64_t i = 0; i < r; i++)
	{
		const auto& h = seeds[i];
		int64_t* o = out + 10 * i;
		o[0] = h.nodeID;
		o[1] = h.nodeOffset;
		o[2] = h.seqPos;
		o[3] = h.matchLen;
		o[4] = h.reverse;
		o[5] = h.alignmentGraphNodeId;
		o[6] = h.alignmentGraphNodeOffset;
		o[7] = (int64_t)h.rawSeedGoodness;
		o[8] = (int64_t)h.seedGoodness;
		o[9] = (int64_t)h.seedClusterSize;
	}
	return 0;
}

// libstdSort (or, with heapOnly, its heap sort) against libstdc++'s std::sort
// (std::partial_sort over the whole range) on n random keys below tieRange:
// 0 when the permutations (original indices) are identical.
int pipe_sort_check(int64_t n, uint64_t seed, uint64_t tieRange, int heapOnly)
{
	std::mt19937_64 rng(seed);
	std::vector<std::pair<uint64_t, uint32_t>> a(n);
	for (int64_t i = 0; i < n; i++) a[i] = {rng() % tieRange, (uint32_t)i};
	auto b = a;
	auto less = [](const std::pair<uint64_t, uint32_t>& l, const std::pair<uint64_t, uint32_t>& r) { return l.first < r.first; };
	if (heapOnly)
	{
		std::partial_sort(a.begin(), a.end(), a.end(), less);
		gad::heapSortRange(b.data(), (int64_t)n, less);
	}
	else
	{
		std::sort(a.begin(), a.end(), less);
		gad::SortRange stack[gad::kSortStackEntries];
		gad::libstdSort(b.data(), (int64_t)n, less, stack);
	}
	for (int64_t i = 0; i < n; i++)
		if (a[i].second != b[i].second) return 1;
	return 0;
}

// Whole pipeline; arguments follow the oracle's orc_align_reads.
struct PipeOptions
{
	int32_t bandwidth;
	int32_t errorCost;
	int32_t xdrop;
	int32_t sloppy;
	int64_t maxCellsPerSlice;
	double seedExtendDensity;
	double multimapScoreFraction;
	double seedDensity;
	int64_t minClusterSize;
	int32_t deviceLoop;
	int32_t maxSpeculativeSeeds;
	double minAlignmentScore;
	int32_t deviceSeedOrder;
	int32_t cigarMatchMismatchMerge;
};

void pipe_set_fault(int f) { g_fault = f; }

int pipe_align_reads(void* ih, const PipeOptions* po, int64_t nreads, const char* const* names, const char* const* seqs, const int64_t* lens, char** gafOut, int32_t* statusOut, int64_t* statsOut)
{
	auto* x = (EmuIndex*)ih;
	ga::AlignOptions o;
	o.ext.bandwidth = po->bandwidth;
	o.ext.errorCost = po->errorCost;
	o.ext.xdrop = po->xdrop;
	o.ext.pad = 0;
	o.ext.maxCellsPerSlice = po->maxCellsPerSlice <= 0 ? INT64_MAX : po->maxCellsPerSlice;
	o.xscoreErrorCost = po->errorCost;
	o.seedExtendDensity = po->seedExtendDensity;
	o.multimapScoreFraction = po->multimapScoreFraction;
	o.minimizerSeedDensity = po->seedDensity;
	o.sloppyOptimizations = po->sloppy != 0;
	o.minSeedClusterSize = po->minClusterSize;
	o.hostThreads = 4;
	o.deviceSeedLoop = po->deviceLoop != 0;
	o.deviceSeedOrder = po->deviceSeedOrder != 0;
	o.cigarMatchMismatchMerge = po->cigarMatchMismatchMerge != 0;
	o.batchedDeviceLoop = po->deviceLoop == 2;
	o.maxSpeculativeSeeds = po->maxSpeculativeSeeds < 1 ? 1 : (uint32_t)po->maxSpeculativeSeeds;
	o.minAlignmentScore = po->minAlignmentScore;
	if (g_proposalOn) o.proposal = &g_proposal;
	if (g_records) o.records = g_records.get();
	std::vector<const char*> nm(names, names + nreads), sq(seqs, seqs + nreads);
	std::vector<uint32_t> ln(nreads);
	for (int64_t i = 0; i < nreads; i++) ln[i] = (uint32_t)lens[i];
	HostBackend be;
	std::vector<ga::ReadOutcome> out;
	ga::AlignStats st;
	try
	{
		ga::alignReads(be, *x->g, x->idx, o, nm, sq, ln, out, st);
		g_lastStats = st;
	}
	catch (const std::exception&)
	{
		return 1;
	}
	size_t total = 1;
	for (auto& r : out) total += r.gaf.size();
	char* buf = (char*)std::malloc(total);
	size_t pos = 0;
	for (int64_t i = 0; i < nreads; i++)
	{
		std::memcpy(buf + pos, out[i].gaf.data(), out[i].gaf.size());
		pos += out[i].gaf.size();
		statusOut[i] = out[i].status;
	}
	buf[pos] = 0;
	*gafOut = buf;
	if (statsOut)
	{
		statsOut[0] = st.rounds;
		statsOut[1] = st.extensions;
		statsOut[2] = st.seeds;
		statsOut[3] = st.wordSteps;
		statsOut[4] = st.unusedExtensions;
		statsOut[5] = st.launches;
	}
	return 0;
}

// AlignOneWay with the caller's seeds for one read; output text in the layout
// of the oracle's orc_align_one_way.
int pipe_align_one_way(void* ih, const PipeOptions* po, const char* seq, int64_t len, const int64_t* seeds, int64_t nseeds, int select, int mapq, char** outText)
{
	auto* x = (EmuIndex*)ih;
	ga::AlignOptions o;
	o.ext.bandwidth = po->bandwidth;
	o.ext.errorCost = po->errorCost;
	o.ext.xdrop = po->xdrop;
	o.ext.pad = 0;
	o.ext.maxCellsPerSlice = po->maxCellsPerSlice <= 0 ? INT64_MAX : po->maxCellsPerSlice;
	o.xscoreErrorCost = po->errorCost;
	o.seedExtendDensity = po->seedExtendDensity;
	o.multimapScoreFraction = po->multimapScoreFraction;
	o.sloppyOptimizations = po->sloppy != 0;
	o.minSeedClusterSize = po->minClusterSize;
	o.hostThreads = 4;
	o.maxSpeculativeSeeds = po->maxSpeculativeSeeds < 1 ? 1 : (uint32_t)po->maxSpeculativeSeeds;
	o.minAlignmentScore = po->minAlignmentScore;
	std::vector<ga::SeedHit> hits(nseeds);
	for (int64_t i = 0; i < nseeds; i++)
	{
		const int64_t* v = seeds + 10 * i;
		hits[i].nodeID = (int)v[0];
		hits[i].nodeOffset = v[1];
		hits[i].seqPos = v[2];
		hits[i].matchLen = v[3];
		hits[i].reverse = v[4] != 0;
		hits[i].alignmentGraphNodeId = v[5] < 0 ? ~0ull : (uint64_t)v[5];
		hits[i].alignmentGraphNodeOffset = v[6] < 0 ? ~0ull : (uint64_t)v[6];
		hits[i].rawSeedGoodness = v[7];
		hits[i].seedGoodness = v[8];
		hits[i].seedClusterSize = v[9];
	}
	HostBackend be;
	std::vector<ga::Alignment> alns;
	ga::ReadOutcome out;
	ga::AlignStats st;
	try
	{
		be.uploadReads({seq}, {(uint32_t)len});
		ga::alignOneWayResident(be, *x->g, o, 0, 1, &hits, &alns, &out, st);
	}
	catch (const std::exception&)
	{
		return -1;
	}
	int status = out.status;
	if (status == 0 && mapq && !ga::addMappingQualities(alns)) status = 2;
	if (status == 0 && select)
	{
		if (!alns.empty()) alns = ga::selectAlignments(alns, o.minAlignmentScore, o.multimapScoreFraction);
		ga::sortForOutput(alns);
	}
	std::string text;
	char buf[256];
	std::snprintf(buf, sizeof(buf), "S %d %llu\n", status, status == 0 ? (unsigned long long)out.seedsExtended : 0ull);
	text += buf;
	if (status == 0)
		for (const auto& a : alns)
		{
			std::snprintf(buf, sizeof(buf), "A %llu %llu %llu %.17g %llu %d %zu\n", (unsigned long long)a.alignmentStart, (unsigned long long)a.alignmentEnd, (unsigned long long)a.alignmentScore, a.alignmentXScore, (unsigned long long)a.seedGoodness, a.mappingQuality, a.trace.size());
			text += buf;
			for (size_t k = 0; k < a.trace.size(); k++)
			{
				const auto& t = a.trace[k];
				std::snprintf(buf, sizeof(buf), "%s%u,%u,%u,%d,%d,%d", k ? " " : "", t.node, t.nodeOffset, t.seqPos, (int)t.nodeSwitch, (int)(unsigned char)t.seqChar, (int)(unsigned char)t.graphChar);
				text += buf;
			}
			text += "\n";
		}
	*outText = (char*)std::malloc(text.size() + 1);
	std::memcpy(*outText, text.c_str(), text.size() + 1);
	return status;
}

void pipe_free(void* p) { std::free(p); }

// proposal with a logistic scorer: w[48], bias; threshold and keepTop as the
// product's ga_context_set_proposal; on = 0 turns it off
void pipe_set_linear_proposal(int on, const float* w, float bias, double threshold, int keepTop, int maxClusters)
{
	g_proposal.maxClusters = maxClusters < 0 ? 0 : (uint32_t)maxClusters;
	g_proposalOn = on != 0;
	if (!g_proposalOn) return;
	g_scorer.w.assign(w, w + ga::kClusterFeatures);
	g_scorer.bias = bias;
	g_proposal.scorer = &g_scorer;
	g_proposal.threshold = threshold;
	g_proposal.keepTop = (uint32_t)keepTop;
}

// outcome records of the following pipe_align_reads calls (on = 1 starts empty)
void pipe_records(int on)
{
	if (on) g_records = std::make_unique<ga::ClusterRecords>();
	else g_records.reset();
}

int64_t pipe_records_take(float** features, float** labels, int64_t** reads)
{
	if (!g_records) return -1;
	g_records->sort();
	const size_t n = g_records->reads.size();
	*features = (float*)std::malloc(std::max<size_t>(1, g_records->features.size()) * sizeof(float));
	*labels = (float*)std::malloc(std::max<size_t>(1, g_records->labels.size()) * sizeof(float));
	*reads = (int64_t*)std::malloc(std::max<size_t>(1, n) * sizeof(int64_t));
	std::memcpy(*features, g_records->features.data(), g_records->features.size() * sizeof(float));
	std::memcpy(*labels, g_records->labels.data(), g_records->labels.size() * sizeof(float));
	std::memcpy(*reads, g_records->reads.data(), n * sizeof(int64_t));
	g_records = std::make_unique<ga::ClusterRecords>();
	return (int64_t)n;
}

// proposal counters of the last pipe_align_reads: clusters scored, clusters and seeds removed
void pipe_proposal_stats(int64_t* out)
{
	out[0] = (int64_t)g_lastStats.proposalClusters;
	out[1] = (int64_t)g_lastStats.prunedClusters;
	out[2] = (int64_t)g_lastStats.prunedSeeds;
}

namespace {
// GraphAligner::exactAlignmentPart (src/GraphAligner.h:403-457) as the reference
// writes it, with its assertions: 0 or 1, or 2 when an assertion fires
int exactPartBisection(const std::vector<ga::TItem>& trace, const ga::SeedHit& seed)
{
	if (!(trace.back().seqPos > trace[0].seqPos)) return 2;
	if (trace.back().seqPos < seed.seqPos) return 0;
	if (trace[0].seqPos > seed.seqPos) return 0;
	size_t high = trace.size();
	size_t low = 0;
	size_t mid = (seed.seqPos - trace[0].seqPos) / (trace.back().seqPos - trace[0].seqPos);
	while (true)
	{
		if (mid >= trace.size()) return 2;
		if (trace[mid].seqPos == seed.seqPos) break;
		if (trace[mid].seqPos < seed.seqPos)
		{
			low = mid;
			mid = (high + low) / 2;
			if (mid == low) mid += 1;
		}
		if (mid >= trace.size()) return 2;
		if (trace[mid].seqPos > seed.seqPos)
		{
			high = mid;
			mid = (high + low) / 2;
		}
		if (!(low < mid) || !(mid < high)) return 2;
	}
	const uint64_t compareNode = (uint64_t)(int64_t)(seed.nodeID * 2 + (seed.reverse ? 1 : 0));
	size_t down = mid;
	while (trace[down].seqPos == seed.seqPos)
	{
		if (compareNode == trace[down].node && seed.nodeOffset == trace[down].nodeOffset) return 1;
		if (down == 0) break;
		down -= 1;
	}
	size_t up = mid;
	while (trace[up].seqPos == seed.seqPos)
	{
		if (compareNode == trace[up].node && seed.nodeOffset == trace[up].nodeOffset) return 1;
		up += 1;
		if (up == trace.size()) break;
	}
	return 0;
}
} // namespace

// The product's exactAlignmentPart against the reference's bisection above on
// random traces whose read positions rise by zero or one per item (as the trace
// rebuild makes them), seeds on and off the traces: the number of disagreements
// and, in *bisectionAsserts, how often the bisection's assertions fired.
int64_t pipe_exact_part_check(uint64_t seed, int64_t trials, int64_t* bisectionAsserts, int64_t* found)
{
	std::mt19937_64 rng(seed);
	int64_t bad = 0, asserts = 0, hits = 0;
	for (int64_t t = 0; t < trials; t++)
	{
		ga::Alignment a;
		// every tenth trace: one to three items on one read position, which the
		// reference's first assertion rejects
		const bool flat = t % 10 == 0;
		const size_t n = flat ? 1 + rng() % 3 : 1 + rng() % 3000;
		const uint32_t stillEvery = flat ? 1 : 2 + (uint32_t)(rng() % 12);
		uint32_t pos = (uint32_t)(rng() % 1000);
		for (size_t i = 0; i < n; i++)
		{
			ga::TItem it;
			it.node = (uint32_t)(rng() % 6);
			it.nodeOffset = (uint32_t)(rng() % 4);
			it.seqPos = pos;
			it.nodeSwitch = false;
			it.seqChar = 'A';
			it.graphChar = 'A';
			a.trace.push_back(it);
			if (rng() % stillEvery != 0) pos++;
		}
		const uint64_t front = a.trace[0].seqPos, back = a.trace.back().seqPos;
		for (int q = 0; q < 40; q++)
		{
			ga::SeedHit s;
			s.seqPos = front + (rng() % (back - front + 7)) - 3;
			if (s.seqPos > back + 3) s.seqPos = front;
			uint64_t node = rng() % 6, off = rng() % 4;
			if (rng() % 2)
			{
				// an item at the seed's position, when there is one
				for (const auto& it : a.trace)
					if (it.seqPos == s.seqPos)
					{
						node = it.node;
						off = it.nodeOffset;
						if (rng() % 3 == 0) break;
					}
			}
			s.nodeID = (int)(node / 2);
			s.reverse = node % 2 == 1;
			s.nodeOffset = off;
			const int want = exactPartBisection(a.trace, s);
			int got;
			try
			{
				got = ga::exactAlignmentPart(a, s) ? 1 : 0;
			}
			catch (...)
			{
				got = 2;
			}
			if (got != want) bad++;
			if (want == 2) asserts++;
			if (want == 1) hits++;
		}
	}
	if (bisectionAsserts) *bisectionAsserts = asserts;
	if (found) *found = hits;
	return bad;
}

// host phase times of the last pipe_align_reads (ms): seeding, seed picking,
// trace rebuild and loop replay, selection and GAF, all host work, whole call
void pipe_host_phase_ms(double* out)
{
	out[0] = g_lastStats.seedingMs;
	out[1] = g_lastStats.pickMs;
	out[2] = g_lastStats.replayMs;
	out[3] = g_lastStats.gafMs;
	out[4] = g_lastStats.hostMs;
	out[5] = g_lastStats.totalMs;
}

// clusters of ordered seeds (rows of 10 as pipe_get_seeds) and their features
int64_t pipe_cluster_features(void* ih, const char* seq, int64_t len, const int64_t* seeds, int64_t nseeds, uint32_t* clusterOf, float* features, int64_t cap)
{
	auto* x = (EmuIndex*)ih;
	std::vector<ga::SeedHit> hits(nseeds);
	for (int64_t i = 0; i < nseeds; i++)
	{
		const int64_t* v = seeds + 10 * i;
		hits[i].nodeID = (int)v[0];
		hits[i].nodeOffset = v[1];
		hits[i].seqPos = v[2];
		hits[i].matchLen = v[3];
		hits[i].reverse = v[4] != 0;
		hits[i].alignmentGraphNodeId = v[5] < 0 ? ~0ull : (uint64_t)v[5];
		hits[i].alignmentGraphNodeOffset = v[6] < 0 ? ~0ull : (uint64_t)v[6];
		hits[i].rawSeedGoodness = v[7];
		hits[i].seedGoodness = v[8];
		hits[i].seedClusterSize = v[9];
	}
	std::vector<uint32_t> c;
	const uint32_t n = ga::seedClusters(*x->g, hits, c);
	for (int64_t i = 0; i < nseeds; i++) clusterOf[i] = c[i];
	if ((int64_t)n > cap) return -(int64_t)n;
	ga::clusterFeatures(*x->g, x->idx.maxCount, seq, (uint32_t)len, hits, c, n, n, features);
	return n;
}
}
