// Plain-old-data types shared by the host code and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace gad {

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
	uint32_t numNodes;
	uint32_t firstAmbiguous;
};

struct ExtParams
{
	int32_t bandwidth;
	int32_t errorCost;
	int32_t xdrop;
	int32_t pad;
	int64_t maxCellsPerSlice; // INT64_MAX when unlimited
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

} // namespace gad
