// Whole-read alignment on one GPU: device seed lookup, host seed ordering, the
// seed-extension loop run as rounds of batched device extensions (every active
// read contributes the extensions of its next seed to one launch), trace
// fix-up, alignment selection and GAF text.
//
// Each read's seed loop makes exactly the decisions of GraphAligner::AlignOneWay
// (src/GraphAligner.h:94-181); batching only interleaves independent reads.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "ga_graph.hpp"
#include "ga_index.hpp"
#include "ga_backend.hpp"

namespace ga {

// src/GraphAlignerWrapper.h:11-50 (SeedHit)
struct SeedHit
{
	int nodeID = 0;                 // bigraph node id / 2
	uint64_t nodeOffset = 0;        // offset in the bigraph node
	uint64_t seqPos = 0;            // read position of the seed's last base
	uint64_t matchLen = 0;
	bool reverse = false;
	uint64_t alignmentGraphNodeId = ~0ull;   // ~0: not set (src/GraphAlignerWrapper.h:16-27)
	uint64_t alignmentGraphNodeOffset = ~0ull;
	uint64_t rawSeedGoodness = 0;
	uint64_t seedGoodness = 0;
	uint64_t seedClusterSize = 0;
};

// One trace item (src/GraphAlignerCommon.h:110-151). node/nodeOffset are split
// node and offset before the fix-ups, bigraph id and offset after them. 16 bytes:
// bigraph ids, node offsets and read positions fit 32 bits (reads are at most
// 16 Mbp); every alignment's trace is written once and read by the seed loop
// and the GAF writer, so its size is memory traffic and page faults.
struct TItem
{
	uint32_t node;
	uint32_t nodeOffset;
	uint32_t seqPos;
	bool nodeSwitch;
	char seqChar;
	char graphChar;
	// left uninitialised: traces are sized first and then filled item by item
	TItem() {}
};
static_assert(sizeof(TItem) == 16, "trace items are 16 bytes");

// AlignmentResult::AlignmentItem (src/GraphAlignerCommon.h:289-340) on the
// seed-extension path: the merged, fixed-up trace and its scores.
struct Alignment
{
	std::vector<TItem> trace;
	int32_t traceScore = 0;
	uint64_t seedGoodness = 0;
	uint64_t alignmentStart = 0;
	uint64_t alignmentEnd = 0;
	uint64_t alignmentScore = ~0ull;
	double alignmentXScore = -1;
	int mappingQuality = 255;
	uint32_t seedIndex = ~0u;  // the seed (in the read's loop order) it was extended from
	bool alignmentFailed() const { return alignmentEnd == alignmentStart; }
	uint64_t alignmentLength() const { return alignmentEnd - alignmentStart; }
};

// GraphAligner::exactAlignmentPart (src/GraphAligner.h:403-457) on a trace built
// by the rebuild (read positions rising by at most one per item): whether it
// passes through the seed's node and offset at the seed's read position. Throws
// the rebuild's read-drop exception when the trace's last position does not
// exceed its first.
bool exactAlignmentPart(const Alignment& aln, const SeedHit& seed);

struct SeedProposal;
struct ClusterRecords;

struct AlignOptions
{
	gad::ExtParams ext;
	int32_t xscoreErrorCost = 294;
	double seedExtendDensity = 0.002;
	double multimapScoreFraction = 0.9;
	double minAlignmentScore = 0;
	double minimizerSeedDensity = 5;
	bool sloppyOptimizations = true;
	uint64_t minSeedClusterSize = 1;
	bool cigarMatchMismatchMerge = false;
	int hostThreads = 16;
	uint32_t readsPerBatch = 1u << 20;
	// false: the seed loops advance in rounds, each round extending the next seed
	// of every unfinished read in one launch (both directions as separate jobs);
	// true: every read's whole seed loop runs on the device in one launch
	bool deviceSeedLoop = false;
	// with deviceSeedLoop: the batched persistent form (a fixed set of workers
	// pulling reads from a device counter, one captured graph per launch shape,
	// worker count autotuned); same results
	bool batchedDeviceLoop = false;
	// seed expansion and OrderSeeds of each batch through the backend
	// (ExtendBackend::orderedSeeds: on the GPU for a device context) instead of
	// host threads; same seeds
	bool deviceSeedOrder = false;
	// host-driven rounds: a read's round r > 0 also extends up to 4^r - 1 of the
	// seeds its loop would try next (at most maxSpeculativeSeeds seeds in all,
	// more than 4^r when few reads remain, see speculationJobBudget);
	// the loop's decisions are then replayed in order, so results are those of
	// one seed at a time and extensions of seeds the replay skips are unused.
	// Fewer rounds means fewer launches waiting on their longest extension.
	uint32_t maxSpeculativeSeeds = 64;
	// and, from the second round on, up to this many extension jobs per launch
	// when few reads remain (roughly the wavefronts one launch keeps resident)
	uint32_t speculationJobBudget = 3072;
	// host-driven rounds: the reads are dealt (longest first) into this many
	// groups, at most the backend's extension lanes, whose seed loops run in
	// separate host threads so one group's host work overlaps another's kernels
	uint32_t pipelineGroups = 2;
	// true (default): groups are length classes with equal shares of the bases;
	// false: the reads are dealt round-robin (longest first) and the groups'
	// first launches are staggered. Either way the groups' launches run one at a
	// time.
	bool groupByLength = true;
	// learned proposal: seed clusters it scores low are removed before the
	// read's loop (none: the plain method)
	const SeedProposal* proposal = nullptr;
	// outcome records of every seed cluster (features and labels) are added here
	ClusterRecords* records = nullptr;
};

enum ReadStatus
{
	READ_ALIGNED = 0,
	READ_NO_SEEDS = 1,
	READ_ASSERT = 2,      // a reference assertion fired; the read is dropped
	READ_NO_ALIGNMENT = 3,
	READ_OVERFLOW = 4     // an extension did not fit the largest device workspace
};

struct ReadOutcome
{
	int status = READ_NO_ALIGNMENT;
	std::string gaf;
	uint64_t seedsExtended = 0;
};

struct AlignStats
{
	double lookupMs = 0;        // device seed lookup (+ reverse complement)
	double extendKernelMs = 0;  // device seed extension
	double hostMs = 0;          // host seeding, fix-up, selection, GAF
	double totalMs = 0;
	uint64_t rounds = 0, launches = 0, extensions = 0, wordSteps = 0, seeds = 0, reruns = 0;
	double lookupCallMs = 0;    // wall time of the seed-lookup calls
	double extendCallMs = 0;    // wall time of the extension calls (kernels, set-up, copies)
	uint64_t unusedExtensions = 0; // speculative extensions the replayed loops skipped
	// parts of hostMs in the host-driven rounds: seeds from matches and their
	// order; picking each round's seeds and building jobs; replaying the loops
	// (trace decode, fix-ups, alignments); selection and GAF text
	double seedingMs = 0, pickMs = 0, replayMs = 0, gafMs = 0;
	// batched persistent form: workers per launch and the time their tuning took
	uint64_t workers = 0;
	double tuneMs = 0;
	// learned proposal: clusters scored, clusters and seeds removed, time taken
	// (features, scoring, pruning)
	uint64_t proposalClusters = 0, prunedClusters = 0, prunedSeeds = 0;
	double proposalMs = 0;
};

// The selection's score threshold: --min-alignment-score is a double but
// SelectionOptions holds it as an int (src/AlignmentSelection.h:21, assigned at
// src/Aligner.cpp:402), so it is truncated toward zero; a value the int cannot
// hold (or NaN) converts to INT32_MIN as the x86-64 conversion instruction
// gives it. SelectAlignments filters only when the threshold is above INT32_MIN
// (src/AlignmentSelection.cpp:36).
inline int32_t selectionScoreThreshold(double v)
{
	if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
	return (int32_t)v;
}

// MinimizerSeeder::getSeeds (src/MinimizerSeeder.cpp:521-555) from the device
// lookup's matches of one read.
std::vector<SeedHit> seedsFromMatches(const GraphHost& g, const MinimizerIndex& idx, const gad::SeedMatch* matches, uint32_t nmatches, uint64_t readLen, double density);

// GraphAligner::orderSeedsByChaining (src/GraphAligner.h:250-312). Returns
// false when one of the reference's assertions there fires (the reference then
// drops the read); the seeds are left partly updated. A seed whose
// alignmentGraphNodeId is ~0 (not from the minimizer seeder) is placed by its
// bigraph node and offset.
bool orderSeedsByChaining(const GraphHost& g, std::vector<SeedHit>& seeds);

// Seeds of every read as the reference's seeder + OrderSeeds produce them;
// status (optional) is READ_ASSERT for a read whose OrderSeeds fired a reference
// assertion, else 0.
// order: 0 leaves the seeds in expansion order, 1 runs OrderSeeds on the host,
// 2 runs expansion and OrderSeeds through the backend (ExtendBackend::orderedSeeds)
void seedReads(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, double density, int order, const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens, std::vector<std::vector<SeedHit>>& out, double* lookupMs, std::vector<int32_t>* status = nullptr);

// src/Aligner.cpp:390-606 for resident reads [first, first + count) of the
// backend (one GAF line per selected alignment, '\n'-terminated, in out[i].gaf;
// names[i] names read first + i).
void alignResident(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, const AlignOptions& opt, const char* const* names, uint32_t first, uint32_t count, ReadOutcome* out, AlignStats& stats);

// GraphAligner::AlignOneWay (src/GraphAligner.h:94-181) with the caller's
// ordered seeds, for resident reads [first, first + count): seeds[i] are read
// first + i's seeds; alignments[i] receives its AlignmentResult::alignments in
// the order AlignOneWay leaves them (unselected), out[i].status READ_ALIGNED
// (whether or not any alignment was made), READ_ASSERT (a reference assertion
// fired: no seeds, a seed past the read's end or a check on the way) or
// READ_OVERFLOW, out[i].seedsExtended AlignmentResult::seedsExtended. Runs the
// host-driven rounds whatever opt.deviceSeedLoop says.
void alignOneWayResident(ExtendBackend& ctx, const GraphHost& g, const AlignOptions& opt, uint32_t first, uint32_t count, const std::vector<SeedHit>* seeds, std::vector<Alignment>* alignments, ReadOutcome* out, AlignStats& stats);

// AlignmentSelection::SelectAlignments (src/AlignmentSelection.cpp:31-95) with
// --min-alignment-score and --multimap-score-fraction.
std::vector<Alignment> selectAlignments(std::vector<Alignment>& all, double minScore, double fraction);

// The output order of src/Aligner.cpp:534,549: by start, then by X-score, best first.
void sortForOutput(std::vector<Alignment>& alns);

// AlignmentSelection::AddMappingQualities (src/AlignmentSelection.cpp:97-135);
// false when one of its assertions fires.
bool addMappingQualities(std::vector<Alignment>& alns);

// GraphAligner::AddGAFLine (src/GraphAlignerGAFAlignment.h:37-250): appends the
// line (no newline) to s; false when a reference assertion fires (a letter the
// reference's character tables reject), s then unchanged.
bool gafLine(std::string& s, const GraphHost& g, const char* name, uint64_t readLen, const Alignment& a, bool cigarMatchMismatchMerge);

// Uploads the reads, then alignResident over all of them.
void alignReads(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, const AlignOptions& opt, const std::vector<const char*>& names, const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens, std::vector<ReadOutcome>& out, AlignStats& stats);

} // namespace ga
