/*
 * graphaligner_amd: C ABI of the MI355X seed-extension path of GraphAligner.
 *
 * Plain C types only. Every call returns a ga_status_t. Handles are opaque.
 * Graphs and indexes are immutable after construction and may be shared
 * between threads; a ga_context_t (one GPU, one stream) is not thread-safe.
 * All GPU work needs a gfx950 device: without one, ga_context_create fails with
 * GA_ENODEVICE and nothing falls back to the CPU.
 */
#ifndef GRAPHALIGNER_AMD_H
#define GRAPHALIGNER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
	GA_OK = 0,
	GA_EXT_FAILED = 1,  /* extension kept no slice: OnewayTrace::TraceFailed (src/GraphAlignerCommon.h:154-163) */
	GA_ASSERT = 2,      /* a reference assertion fires; the reference drops the read (src/Aligner.cpp:493-500) */
	GA_EOVERFLOW = 3,   /* device workspace too small even after reruns */
	GA_EBADARG = 4,
	GA_EIO = 5,
	GA_ENODEVICE = 6,
	GA_EHIP = 7,
	GA_ENOMEM = 8,
	GA_ESTALE = 9       /* ga_align_resident after another call replaced the uploaded read set */
} ga_status_t;

typedef struct ga_graph ga_graph_t;
typedef struct ga_index ga_index_t;
typedef struct ga_context ga_context_t;

/* Aligner parameters: the AlignerParams fields on the seed-extension path
 * (src/Aligner.h:10-49, defaults src/AlignerMain.cpp:119-152, presets :157-183). */
typedef struct
{
	int64_t bandwidth;               /* -b, alignmentBandwidth */
	int64_t maxCellsPerSlice;        /* -C, tangle effort; <= 0 means unlimited */
	double preciseClippingIdentity;  /* --precise-clipping (0.66) -> XscoreErrorCost */
	int32_t xdrop;                   /* --X-drop; < 0 selects the reference default */
	int32_t tryAllSeeds;             /* --try-all-seeds (disables sloppy optimizations) */
	double seedExtendDensity;        /* --seeds-extend-density; -1 extends all seeds */
	int64_t seedClusterMinSize;      /* --seeds-clustersize */
	double multimapScoreFraction;    /* --multimap-score-fraction */
	double minAlignmentScore;        /* --min-alignment-score */
	int32_t minimizerLength;         /* --seeds-minimizer-length */
	int32_t minimizerWindowSize;     /* --seeds-minimizer-windowsize */
	double minimizerSeedDensity;     /* --seeds-minimizer-density; -1 = all */
	double minimizerDiscardMostNumerousFraction; /* --seeds-minimizer-ignore-frequent */
	int32_t cigarMatchMismatchMerge; /* --cigar-match-mismatch */
	/* Execution only (no reference counterpart, results are identical): 0 runs
	 * the seed loops in rounds, every round extending the next seed of each
	 * unfinished read in one GPU launch; 1 runs each read's whole seed loop on
	 * the GPU, all reads in one launch. */
	int32_t deviceSeedLoop;
	/* Execution only, rounds mode: from a read's second round on, round r also
	 * extends up to 4^r - 1 of the seeds its loop would try next (more when few
	 * reads remain), at most this many seeds in all (1: one seed per round). The
	 * loop's decisions are replayed in order, so results do not depend on it.
	 * Presets set 64. */
	int32_t maxSpeculativeSeeds;
} ga_params_t;

/* Fills params with the defaults plus preset "vg" or "dbg" (-x), as
 * src/AlignerMain.cpp:119-231 does, X-drop default included. */
ga_status_t ga_params_preset(const char* preset, ga_params_t* params);

/* Graph: GfaGraph::LoadFromFile(path, true) + DirectedGraph::BuildFromGFA
 * (src/Aligner.cpp:640-646, src/BigraphToDigraph.cpp:215-267). */
ga_status_t ga_graph_load_gfa(const char* path, ga_graph_t** graph, char* err, size_t errcap);
void ga_graph_free(ga_graph_t* graph);
int64_t ga_graph_num_nodes(const ga_graph_t* graph);
int64_t ga_graph_num_edges(const ga_graph_t* graph);
int64_t ga_graph_first_ambiguous(const ga_graph_t* graph);
/* Copies the per-split-node arrays of AlignmentGraph (src/AlignmentGraph.h:121-139):
 * lengths, bigraph ids, offsets, component numbers, linearizable, reverse,
 * chain numbers and positions, 2-bit chunks (2 per node; zero for IUPAC nodes),
 * and in/out adjacency in CSR form (offsets have num_nodes + 1 entries). */
ga_status_t ga_graph_export(const ga_graph_t* graph, int64_t* nodeLength, int64_t* nodeIds, int64_t* nodeOffset, int64_t* component, uint8_t* linearizable, uint8_t* reverse, int64_t* chain, int64_t* chainPos, uint64_t* chunks, int64_t* inOffsets, int64_t* inNeighbors, int64_t* outOffsets, int64_t* outNeighbors);

/* One GPU (HIP device ordinal), one stream, resident graph copies. */
ga_status_t ga_context_create(int device, ga_context_t** ctx, char* err, size_t errcap);
void ga_context_free(ga_context_t* ctx);
/* Waits until all work on the context's device has finished. */
ga_status_t ga_context_synchronize(ga_context_t* ctx);

/* Per-slice words of one extension's DP table, for parity checks. */
typedef struct
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
	uint64_t itemsDigest;
	int32_t scoresNotValid;
	int32_t pad2;
} ga_slice_summary_t;

/* Batched GraphAlignerBitvectorBanded::getReverseTraceFromSeed
 * (src/GraphAlignerBitvectorBanded.h:59-76): extension i aligns seqs[i][0, lens[i])
 * from offset offsets[i] of bigraph node bigraphNodeIds[i]. Trace i is written to
 * traces + 4 * i * traceCap as rows (split node, node offset, seqPos, nodeSwitch),
 * end first, exactly as OnewayTrace::trace before reversal; traceCap >= 2*len+256.
 * status[i] is GA_OK, GA_EXT_FAILED or GA_ASSERT. Summaries (optional, may be NULL)
 * take sumCap rows per extension. */
ga_status_t ga_extend_batch(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs);

/* Minimizer index over the graph: MinimizerSeeder's constructor and initMaxCount
 * (src/MinimizerSeeder.cpp:299-491, 557-575) with k = params->minimizerLength,
 * window = params->minimizerWindowSize and the most numerous
 * params->minimizerDiscardMostNumerousFraction of k-mers left out of seeding.
 * Built on the host; the table is copied to a device on its first use there. */
ga_status_t ga_index_build(const ga_graph_t* graph, const ga_params_t* params, ga_index_t** index, char* err, size_t errcap);
void ga_index_free(ga_index_t* index);
int64_t ga_index_max_count(const ga_index_t* index);
int64_t ga_index_num_kmers(const ga_index_t* index);

/* Seeds of n reads: MinimizerSeeder::getSeeds (src/MinimizerSeeder.cpp:521-555)
 * with the k-mer lookup on the GPU and, when order != 0,
 * GraphAligner::orderSeedsByChaining (src/GraphAligner.h:250-312). *seeds is a
 * buffer from malloc (free with ga_free) holding counts[0] + ... + counts[n-1]
 * seeds, read by read, 10 int64 each in SeedHit field order
 * (src/GraphAlignerWrapper.h:11-50): nodeID, nodeOffset, seqPos, matchLen,
 * reverse, alignmentGraphNodeId, alignmentGraphNodeOffset, rawSeedGoodness,
 * seedGoodness, seedClusterSize. */
ga_status_t ga_seeds(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, double density, int32_t order, int64_t n, const char* const* seqs, const int64_t* lens, int64_t** seeds, int64_t* counts);

typedef struct
{
	double lookupMs;       /* device seed lookup and reverse complements */
	double extendKernelMs; /* device seed extension, summed over launches */
	double hostMs;         /* host seed ordering, trace fix-up, selection, GAF text */
	double totalMs;
	int64_t rounds;        /* device seed-loop launches, workspace-overflow reruns included */
	int64_t launches;
	int64_t extensions;    /* one-direction seed extensions run */
	int64_t wordSteps;     /* forward 64-row column steps of the bit-parallel DP */
	int64_t seeds;         /* seeds found (before the seed loop's skips) */
	int64_t reruns;        /* reads rerun with larger device workspaces */
	double lookupCallMs;   /* wall time of the seed-lookup calls: kernels, copies and match gathering */
	double extendCallMs;   /* wall time of the extension calls: kernels, workspace set-up and copies */
	int64_t unusedExtensions; /* extensions run ahead for later seeds that the seed loop then skipped */
	double seedingMs;      /* parts of hostMs (rounds mode): seeds from matches and their order, */
	double pickMs;         /* picking each round's seeds, */
	double replayMs;       /* replaying the seed loops on the results (trace fix-ups, alignments), */
	double gafMs;          /* selection and GAF text */
	int64_t gafBytes;      /* length of the returned GAF text (without its NUL) */
} ga_align_stats_t;

/* Per-read alignment of src/Aligner.cpp:390-606 (seed, order, extend in a seed
 * loop, select, GAF) for n reads with the given parameters. The seed loop of
 * each read makes the decisions of GraphAligner::AlignOneWay
 * (src/GraphAligner.h:94-181); every read's whole seed loop runs on the GPU
 * (one wavefront per read, all reads in one launch). *gaf is a NUL-terminated buffer from malloc (free with
 * ga_free): the GAF lines (src/GraphAlignerGAFAlignment.h:37-250) of read 0,
 * then read 1, ... status[i]: 0 aligned, 1 no seeds, 2 a reference assertion
 * fired and the read was dropped, 3 no alignment kept, 4 an extension exceeded
 * the largest device workspace. stats may be NULL. */
ga_status_t ga_align_reads(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t n, const char* const* names, const char* const* seqs, const int64_t* lens, char** gaf, int32_t* status, ga_align_stats_t* stats);

/* Reads resident on the context's GPU: ga_reads_upload copies n reads into HBM
 * (and keeps a host copy for trace fix-ups and GAF text), replacing the previous
 * set; ga_align_resident then runs ga_align_reads' per-read pipeline on resident
 * reads [first, first + count) with no host-to-device read transfer. names[i]
 * names read first + i; gaf and status are as in ga_align_reads. */
ga_status_t ga_reads_upload(ga_context_t* ctx, int64_t n, const char* const* seqs, const int64_t* lens);
ga_status_t ga_align_resident(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t first, int64_t count, const char* const* names, char** gaf, int32_t* status, ga_align_stats_t* stats);

void ga_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
