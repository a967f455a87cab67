// What the alignment pipeline needs from a device: resident reads with their
// index matches, and batched seed extensions over them. Context (ga_runtime.hpp)
// is the gfx950 implementation; this header has no HIP dependency.
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>
#include "ga_device_types.h"
#include "ga_trace_codes.h"
#include "ga_graph.hpp"
#include "ga_index.hpp"
#include "ga_order_core.h"

namespace ga {

struct SeedHit;


// One extension: align seq[0, len) from column startOffset of split node startNode.
// With a resident sequence base (extend's devSeqBase) the sequence is the
// resident one at devOffset and seq is not read.
struct ExtJob
{
	const char* seq;
	uint32_t len;
	uint32_t startNode;
	uint32_t startOffset;
	uint64_t devOffset = 0;
	// optional per-slice bandwidth schedule (host memory, bandLen entries; slice
	// s uses bands[min(s, bandLen - 1)]); bandLen 0 uses the fixed bandwidth
	const int32_t* bands = nullptr;
	uint32_t bandLen = 0;
};

// Traces come back in the compact form of ga_trace_codes.h: one code byte per
// item, plus the escaped items whole.
struct ExtBatch
{
	std::vector<gad::ExtResult> results;
	std::vector<uint64_t> traceOffset; // per job, into codes
	const uint8_t* codes = nullptr;    // owned by the backend; valid until its next call
	std::vector<std::pair<uint64_t, uint64_t>> escapes; // (index into codes, packed item), by index
	std::vector<gad::SliceSummary> summaries;
	std::vector<uint64_t> summaryOffset;
	double kernelMs = 0;               // summed kernel time of all launches
	uint32_t launches = 0;
	uint64_t wordSteps = 0;
	uint64_t reruns = 0;               // jobs rerun with larger workspaces

	// packed items of job j's trace (results[j].ntrace of them, end first)
	void items(const GraphHost& g, size_t j, std::vector<uint64_t>& out) const
	{
		const uint32_t n = results[j].ntrace;
		out.resize(n);
		if (n == 0) return;
		const gad::TraceGraph tg{g.nodeLen.data(), g.inOff.data(), g.inAdj.data()};
		const uint64_t base = traceOffset[j];
		const uint8_t* c = codes + base;
		auto e = std::lower_bound(escapes.begin(), escapes.end(), std::make_pair(base, (uint64_t)0));
		uint64_t prev = 0;
		for (uint32_t k = 0; k < n; k++)
		{
			if (c[k] == gad::kTraceEscape)
			{
				if (e == escapes.end() || e->first != base + k) throw std::runtime_error("trace codes: escaped item missing");
				prev = e->second;
				++e;
			}
			else
			{
				if (k == 0) throw std::runtime_error("trace codes: first item not escaped");
				prev = gad::traceDecode(tg, prev, c[k]);
			}
			out[k] = prev;
		}
	}
};

// Encodes traces[j] (packed items, end first) of every job into out's codes
// (stored in codeStore) and escapes, with out.traceOffset; the host form of the
// device encoder, for backends that compute traces on the host.
inline void encodeTraces(const GraphHost& g, const std::vector<std::vector<uint64_t>>& traces, std::vector<uint8_t>& codeStore, ExtBatch& out)
{
	const gad::TraceGraph tg{g.nodeLen.data(), g.inOff.data(), g.inAdj.data()};
	codeStore.clear();
	out.escapes.clear();
	out.traceOffset.assign(traces.size(), 0);
	for (size_t j = 0; j < traces.size(); j++)
	{
		out.traceOffset[j] = codeStore.size();
		const std::vector<uint64_t>& t = traces[j];
		for (size_t k = 0; k < t.size(); k++)
		{
			uint32_t code = k == 0 ? gad::kTraceEscape : gad::traceCode(tg, t[k - 1], t[k]);
			if (code == gad::kTraceEscape) out.escapes.emplace_back(codeStore.size(), t[k]);
			codeStore.push_back((uint8_t)code);
		}
	}
	out.codes = codeStore.data();
}

// Reads resident on the device: forward bytes (and, once looked up, their
// reverse complements) plus a host copy for the trace fix-ups and GAF text.
struct ReadSet
{
	std::vector<uint64_t> offset;  // per read, into the forward bytes
	std::vector<uint32_t> length;
	std::vector<char> host;        // forward bytes
	uint64_t bytes = 0;
	uint32_t maxLen = 0;
	size_t size() const { return offset.size(); }
	const char* seq(size_t i) const { return host.data() + offset[i]; }
};

// Index matches of a range of resident reads: matches of read first + i are
// matches[matchOffset[i], matchOffset[i] + matchCount[i]).
struct SeedMatches
{
	std::vector<gad::SeedMatch> matches;
	// when set, the matches are here instead (a backend's staging memory, valid
	// until its next lookup) and `matches` is empty
	const gad::SeedMatch* staged = nullptr;
	std::vector<uint64_t> matchOffset;
	std::vector<uint32_t> matchCount;
	double lookupMs = 0;
	const gad::SeedMatch* data() const { return staged ? staged : matches.data(); }
};

// Workspace capacities of the first extension attempt; a job that overflows is
// rerun with every capacity multiplied by kArenaGrowth, at most kArenaAttempts
// attempts in all. Band items are budgeted per 64-row slice.
constexpr int kArenaAttempts = 4;
constexpr uint32_t kArenaGrowth = 8;
inline gad::ArenaCaps initialArenaCaps()
{
	gad::ArenaCaps c;
	c.maxSlices = 0;  // set per request from its length
	c.maxItems = 48;
	// the band queue of a first attempt fits the single-lane kernel's LDS copy;
	// reruns grow past it
	c.maxSlots = gad::kLdsSlots;
	c.maxExtras = gad::kLdsExtras;
	c.maxPending = gad::kLdsPending;
	c.pad = 0;
	return c;
}
inline void growArenaCaps(gad::ArenaCaps& c)
{
	c.maxItems *= kArenaGrowth;
	c.maxSlots *= kArenaGrowth;
	c.maxExtras *= kArenaGrowth;
	c.maxPending *= kArenaGrowth;
}

// Trace capacity of an extension of len read bases (a trace has at most one item
// per read base plus one per graph base).
inline uint32_t traceCapFor(uint32_t len) { return 2 * len + 256; }

// One read for the device seed loop: its resident index, its ordered seeds
// (seeds[seedStart, seedStart + seedCount)) and its extension budget.
struct ReadLoopJob
{
	uint32_t read;
	uint32_t seedCount;
	uint64_t seedStart;
	uint64_t extendSeeds;
};

struct ReadLoopOutput
{
	std::vector<gad::ReadResult> results;   // per job
	std::vector<uint64_t> alnOffset;         // per job, into alns
	std::vector<gad::DevAlignment> alns;     // creation order per job
	std::vector<uint64_t> poolOffset;        // per job, into pool
	std::vector<uint64_t> pool;              // raw trace items
	double kernelMs = 0;
	uint32_t launches = 0;
	uint64_t wordSteps = 0, extensions = 0;
	uint64_t reruns = 0;                     // jobs rerun with larger workspaces
	uint32_t workers = 0;                    // batched persistent form: workers per launch
	double tuneMs = 0;                       // and the time its worker-count tuning took
};

// Workspace capacities of the device seed loop; grown like the arena on overflow.
struct ReadLoopCaps
{
	gad::ArenaCaps arena;
	uint32_t poolPerBase;  // alignment trace pool items per read base
	uint32_t poolExtra;
	uint32_t alnCap;       // alignment records per read
};
inline ReadLoopCaps initialReadLoopCaps()
{
	ReadLoopCaps c;
	c.arena = initialArenaCaps();
	// a read's whole seed loop shares one workspace, and an overflow reruns the
	// whole read: start wider than a single extension's first attempt
	c.arena.maxItems *= 2;
	c.arena.maxSlots = 1024;
	c.arena.maxExtras = 8192;
	c.arena.maxPending = 1024;
	c.poolPerBase = 12;
	c.poolExtra = 4096;
	c.alnCap = 64;
	return c;
}
inline void growReadLoopCaps(ReadLoopCaps& c)
{
	growArenaCaps(c.arena);
	c.poolPerBase *= 4;
	c.alnCap *= 4;
}
inline uint32_t poolCapFor(uint32_t len, const ReadLoopCaps& c)
{
	uint64_t v = (uint64_t)len * c.poolPerBase + c.poolExtra;
	return v > 0xffffffffull ? 0xffffffffu : (uint32_t)v;
}

class ExtendBackend
{
public:
	virtual ~ExtendBackend() = default;
	// Extension calls on different lanes (0 .. extendLanes() - 1) may run at the
	// same time from different threads; calls on one lane may not.
	virtual uint32_t extendLanes() const { return 1; }
	// Runs all jobs; statuses EXT_OVERFLOW are rerun with larger workspaces and
	// only remain if the largest workspace still does not fit.
	virtual void extend(const GraphHost& g, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, bool wantSummaries, ExtBatch& out, const char* devSeqBase = nullptr, uint32_t lane = 0) = 0;
	// Makes the reads resident (replacing the previous set).
	virtual void uploadReads(const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens) = 0;
	virtual const ReadSet& reads() const = 0;
	// For resident reads [first, first + count): writes their reverse complements
	// and looks every read k-mer up in the index.
	virtual void lookupSeeds(const MinimizerIndex& idx, uint32_t first, uint32_t count, SeedMatches& out) = 0;
	// The seeds of resident reads [first, first + count) from the matches sm that
	// lookupSeeds just returned for that range: MinimizerSeeder's expansion
	// (src/MinimizerSeeder.cpp:521-555) and OrderSeeds (src/GraphAligner.h:250-312)
	// of every read, run by the backend (ga_order_core.h). seeds[i] receives read
	// i's ordered seeds; asserted[i] = 1 when OrderSeeds fired a reference
	// assertion (the read is dropped).
	virtual void orderedSeeds(const GraphHost& g, const MinimizerIndex& idx, uint32_t first, uint32_t count, const SeedMatches& sm, double density,
	                          std::vector<std::vector<SeedHit>>& seeds, std::vector<uint8_t>& asserted) = 0;
	// For resident reads [first, first + count): writes their reverse complements
	// (what lookupSeeds does besides the lookup), for seeds that come from the caller.
	virtual void prepareReads(uint32_t first, uint32_t count) = 0;
	// Start of an alignment call: per-read device data derived from the resident
	// reads (reverse complements, match-mask planes) is computed again by the call
	// instead of being reused from an earlier one.
	virtual void beginAlignmentCall() {}
	// Forward bytes of resident read i start at readsBase() + offset[i], its reverse
	// complement (once looked up) at readsBase() + reverseBase() + offset[i].
	virtual const char* readsBase() const = 0;
	virtual uint64_t reverseBase() const = 0;
	// Runs the seed loop of every job's read (the reads' reverse complements must
	// be in place, i.e. lookupSeeds ran on them); jobs whose workspaces overflow
	// are rerun with larger ones, kArenaAttempts attempts in all.
	virtual void runReadLoops(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out) = 0;
	// The same read loops in the batched persistent form (same outputs); a
	// backend without one runs runReadLoops.
	virtual void runReadLoopsBatched(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out) { runReadLoops(g, p, seeds, jobs, out); }
};

} // namespace ga
