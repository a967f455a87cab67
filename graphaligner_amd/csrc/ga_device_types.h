// Plain-old-data types shared by the host code and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace gad {

// The static data of one split node in one 48-byte record, so that a node's
// calculation reads it in one memory round trip: its 2-bit bases (zero words
// for IUPAC nodes), length, component, in-neighbour range in inAdj and the first
// in-neighbour (the one a linearizable node's test reads), and out-edge range in
// outEdges.
struct NodeRec
{
	uint64_t seq0, seq1;
	uint32_t len;
	uint32_t comp;
	uint32_t inBegin, inCount;
	uint32_t outBegin, outCount;
	uint32_t firstIn; // inAdj[inBegin], or ~0 without in-neighbours
	uint32_t linearizable;
};
static_assert(sizeof(NodeRec) == 48, "NodeRec is three 16-byte loads");

// An out-edge with its target's component number (the band queue's key), read
// together when a node's changed end relaxes its out-neighbours.
struct OutEdge
{
	uint32_t target;
	uint32_t comp;
};

// Read-only alignment graph resident in HBM. Node sequences are 2-bit packed
// (A0 C1 G2 T3, LSB first, 32 bp per 64-bit chunk); IUPAC nodes (index >=
// firstAmbiguous) carry four 64-bit base masks instead. Adjacency is CSR.
struct DevGraph
{
	const uint32_t* nodeLen;
	const uint64_t* seq2;
	const uint64_t* amb4;
	const uint32_t* comp;
	const uint8_t* linearizable;
	const uint32_t* inOff;
	const uint32_t* inAdj;
	const uint32_t* outOff;
	const uint32_t* outAdj;
	const int32_t* nodeId;   // bigraph node id of each split node
	const uint32_t* nodeOff; // offset of the split node in its bigraph node
	const uint32_t* bigSize; // length of the split node's bigraph node
	const NodeRec* rec;      // per split node, the fields above that the DP reads
	const OutEdge* outEdges; // out-edges in outAdj order, with target components
	uint32_t numNodes;
	uint32_t firstAmbiguous;
};

struct ExtParams
{
	int32_t bandwidth;
	int32_t errorCost;
	int32_t xdrop;
	int32_t pad = 0;
	int64_t maxCellsPerSlice; // INT64_MAX when unlimited
	// per-slice bandwidth schedules of a launch's extensions (ExtRequest::bandOff)
	const int32_t* bandSched = nullptr;
};

// One seed extension in one direction: align seq[0, len) starting right after
// column startOffset of split node startNode (getReverseTraceFromSeed).
struct ExtRequest
{
	uint64_t seqOffset;   // into the sequence buffer
	uint64_t traceOffset; // into the trace buffer (packed items)
	uint32_t len;
	uint32_t startNode;
	uint32_t startOffset;
	uint32_t traceCap;
	// per-slice bandwidth schedule: slice s uses ExtParams::bandSched[bandOff +
	// min(s, bandLen - 1)]; bandLen 0 uses ExtParams::bandwidth for every slice
	uint32_t bandOff = 0;
	uint32_t bandLen = 0;
};

enum ExtStatus : int32_t
{
	EXT_OK = 0,
	EXT_FAILED = 1,   // no slice survived: OnewayTrace::TraceFailed
	EXT_ASSERT = 2,   // a reference assertion would fire: the read is dropped
	EXT_OVERFLOW = 3, // a per-wave workspace capacity was exceeded: rerun larger
};

struct ExtResult
{
	int32_t status;
	int32_t score;
	uint32_t ntrace;
	uint32_t nslices;     // slices kept in the DP table (excluding the initial one)
	uint64_t wordSteps;   // forward Myers column steps (64 cells each)
	uint64_t nodeCalcs;   // forward node calculations
	uint32_t maxBand;     // largest band (nodes) of a kept slice
	uint32_t overflowWhat;
	// device real-time clock (100 MHz) at the extension's start and end, set by
	// the single-lane kernel (zero elsewhere); diagnostics only
	uint64_t clockStart;
	uint64_t clockEnd;
};

// Per-slice record for parity checks (same fields as the oracle's SliceSummary).
struct SliceSummary
{
	int64_t j;
	int32_t minScore;
	int32_t pad0;
	int64_t minScoreNode;
	int64_t minScoreNodeOffset;
	int32_t maxExactEndposScore;
	int32_t pad1;
	int64_t maxExactEndposNode;
	int64_t cellsProcessed;
	int64_t numNodes;
	uint64_t itemsHash;
	int32_t scoresNotValid;
	int32_t pad2;
};

// Capacities of one wave's workspace (all counts, not bytes).
// Band-queue capacities that the single-lane extension kernel holds in LDS
// (powers of two for the slot table); the host's first-attempt capacities
constexpr uint32_t kLdsSlots = 32;
constexpr uint32_t kLdsExtras = 64;
constexpr uint32_t kLdsPending = 32;
// default cap on the single-lane extension kernel's workgroups per SIMD
constexpr uint32_t kSingleLaneWavesPerSimd = 5;

struct ArenaCaps
{
	uint32_t maxSlices;   // slice headers (DP table rows + 1)
	uint32_t maxItems;    // band items over all kept slices
	uint32_t maxSlots;    // nodes touched in one slice
	uint32_t maxExtras;   // queued incoming edges in one slice
	uint32_t maxPending;  // simultaneously queued nodes
	uint32_t pad;
};

// DP-table slice header. j = (index - 1) * 64.
struct SliceHdr
{
	int32_t minScore;
	int32_t minScoreNode;
	int32_t minScoreNodeOffset;
	int32_t maxExactEndposScore;
	int32_t maxExactEndposNode;
	int32_t bandwidth;
	int32_t scoresNotValid;
	uint32_t itemStart;
	uint32_t itemCount;
	uint32_t pad;
	int64_t cells;
};

// Band item: NodeSliceMapItem (start/end column words, bottom-row HP/HN, minScore)
struct Item
{
	uint64_t sVP, sVN, eVP, eVN, HP, HN;
	int32_t sEnd, eEnd, minScore;
	uint32_t node;
};

// Minimizer table slot (k-mer -> position range); kmer == ~0 marks an empty slot.
struct SeedTableSlot
{
	uint64_t kmer;
	uint32_t start;
	uint32_t count;
};

// One read k-mer found in the index: read position of its last base, and the
// k-mer's range in the position array.
struct SeedMatch
{
	uint32_t pos;
	uint32_t start;
	uint32_t count;
};

// One seed of the device seed loop: the SeedHit fields the loop reads
// (src/GraphAligner.h:94-181) plus the split-node start of each extension.
struct DevSeed
{
	uint64_t seedGoodness;
	uint64_t clusterSize;
	uint32_t seqPos;
	int32_t nodeID;      // bigraph node id / 2
	uint32_t nodeOffset; // offset in the bigraph node
	uint32_t reverse;
	uint32_t bwdNode, bwdOff; // backward extension start (split node, offset)
	uint32_t fwdNode, fwdOff; // forward extension start
};

struct ReadLoopParams
{
	ExtParams ext;
	uint64_t minSeedClusterSize;
	uint32_t sloppy;
	uint32_t pad;
};

// One read's seed loop on the device. Offsets locate its bytes, seeds and
// workspaces; capacities are item / record counts.
struct ReadJob
{
	uint64_t fwdOffset;   // forward bytes in the read buffer
	uint64_t revOffset;   // reverse-complement bytes in the read buffer
	uint64_t seedStart;   // into the seed array
	uint64_t arenaOffset; // extension workspace (bytes)
	uint64_t traceOffset; // two raw trace regions of traceCap items each
	uint64_t poolOffset;  // alignment trace pool (items)
	uint64_t alnOffset;   // alignment records
	uint64_t extendSeeds; // seed extension budget (AlignOneWay's extendSeeds)
	uint32_t len;
	uint32_t seedCount;
	uint32_t traceCap;
	uint32_t poolCap;
	uint32_t alnCap;
	uint32_t pad;
};

// One alignment made by the device seed loop, in creation order: its trace is
// the backward extension's raw trace (nb items, end first) followed by the
// forward extension's raw trace (nf items, end first) in the read's pool.
struct DevAlignment
{
	uint64_t seedGoodness;
	uint32_t poolOffset; // relative to the read's pool
	uint32_t seedIndex;  // in the read's ordered seeds
	uint32_t start, end; // read interval [start, end)
	uint32_t nb, nf;
};

enum ReadLoopStatus : int32_t
{
	READLOOP_DONE = 0,
	READLOOP_ASSERT = 2,   // a reference assertion fired: the read is dropped
	READLOOP_OVERFLOW = 4, // a workspace capacity was exceeded: rerun larger
};

struct ReadResult
{
	int32_t status;
	uint32_t nAlignments;
	uint32_t seedsExtended;
	uint32_t poolUsed;
	uint64_t wordSteps;
	uint32_t extensions;
	uint32_t overflowWhat;
};

} // namespace gad
