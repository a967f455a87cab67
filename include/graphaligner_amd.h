/*
 * graphaligner_amd: C ABI of the MI355X seed-extension path of GraphAligner.
 *
 * Plain C types only. Every call returns a ga_status_t. Handles are opaque.
 * Graphs and indexes are immutable after construction and may be shared
 * between threads; a ga_context_t (one GPU) is not thread-safe.
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
	GA_ESTALE = 9,      /* ga_align_resident after another call replaced the uploaded read set */
	GA_ESCRATCH = 10    /* the batch exceeds ga_extend_batch_noalloc's fixed scratch */
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
	 * the GPU, all reads in one launch; 2 runs them in the batched persistent
	 * form: a fixed set of workers (count autotuned on the first batch) pulls
	 * reads, longest first, from a device work counter, launched as one captured
	 * HIP graph per batch. */
	int32_t deviceSeedLoop;
	/* Execution only, rounds mode: from a read's second round on, round r also
	 * extends up to 4^r - 1 of the seeds its loop would try next (more when few
	 * reads remain), at most this many seeds in all (1: one seed per round). The
	 * loop's decisions are replayed in order, so results do not depend on it.
	 * Presets set 64. */
	int32_t maxSpeculativeSeeds;
	/* Execution only: 1 runs seed expansion and OrderSeeds of every read on the
	 * GPU (one thread per read, ga_seeds' order 2) instead of host threads; the
	 * seeds are the same. */
	int32_t deviceSeedOrder;
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

/* One GPU (HIP device ordinal): a stream for uploads and seed lookups and four
 * extension lanes with their own streams, resident graph and index copies. */
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

/* ga_extend_batch with a per-slice bandwidth schedule: the hook through which
 * getXdropSlices picks each slice's bandwidth (src/GraphAlignerBitvectorBanded.h:611,
 * kept in DPSlice.bandwidth, :552, and read back by the traceback's quit scores,
 * src/GraphAlignerBitvectorCommon.h:633,659,684). Slice s of extension i (its
 * rows 64 s .. 64 s + 63) uses schedule[scheduleOffsets[i] + min(s, scheduleLens[i] - 1)];
 * scheduleLens[i] == 0 uses params->bandwidth for all of extension i's slices.
 * Bandwidths must lie in 0 .. 2^20 (else GA_EBADARG). A schedule whose entries
 * all equal params->bandwidth gives exactly ga_extend_batch's results. */
ga_status_t ga_extend_batch_scheduled(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, const int32_t* schedule, const int64_t* scheduleOffsets, const int64_t* scheduleLens, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, ga_slice_summary_t* summaries, int64_t sumCap, int64_t* nslices, double* kernelMs);

/* No counterpart in the reference: ga_extend_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation). Same extensions, trace rows, scores and
 * statuses as ga_extend_batch (no slice summaries); every device buffer of the
 * call (requests, sequences, match masks, traces, workspaces) is carved from
 * one file-scope device array of 128 MiB that each GPU holds until the process
 * ends, so the call makes no hipMalloc/hipFree. Limits: at most 4096
 * extensions, and their sequences, traces (16 * len + 2 KiB each) and
 * workspaces (about 100 bytes per base) within the 128 MiB; a larger batch
 * fails with GA_ESCRATCH and nothing is run. Extensions whose workspaces still
 * overflow after the reruns that fit keep status GA_EOVERFLOW (the call then
 * returns GA_EOVERFLOW). The graph must already be resident on the context's
 * GPU (an earlier ga_extend_batch or alignment call with it made it so), else
 * GA_EBADARG. The scratch is shared by every call on that GPU and keeps the
 * last call's contents: calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialised by the caller. */
ga_status_t ga_extend_batch_noalloc(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const int32_t* bigraphNodeIds, const int64_t* offsets, int64_t traceCap, int64_t* traces, int64_t* ntrace, int32_t* scores, int32_t* status, double* kernelMs);

/* Minimizer index over the graph: MinimizerSeeder's constructor and initMaxCount
 * (src/MinimizerSeeder.cpp:299-491, 557-575) with k = params->minimizerLength,
 * window = params->minimizerWindowSize and the most numerous
 * params->minimizerDiscardMostNumerousFraction of k-mers left out of seeding.
 * Built on the host; the table is copied to a device on its first use there. */
ga_status_t ga_index_build(const ga_graph_t* graph, const ga_params_t* params, ga_index_t** index, char* err, size_t errcap);
/* The same index built on the context's GPU (no reference counterpart of the
 * device build; same constructor and initMaxCount semantics): every position of
 * every bigraph node decides in parallel whether the reference's window walk
 * (iterateMinimizersReal, src/MinimizerSeeder.cpp:104-189) reports it, then the
 * reports are sorted and grouped on the device. Gives ga_index_build's k-mers,
 * position ranges, position order and maxCount (hash-table slots may differ). */
ga_status_t ga_index_build_gpu(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, ga_index_t** index, char* err, size_t errcap);
void ga_index_free(ga_index_t* index);
int64_t ga_index_max_count(const ga_index_t* index);
int64_t ga_index_num_kmers(const ga_index_t* index);
int64_t ga_index_num_positions(const ga_index_t* index);
/* The index's contents for checks: its ga_index_num_kmers k-mers in increasing
 * order with their position ranges (start, count into positions), and the
 * ga_index_num_positions positions (split node << 6 | offset; one k-mer's in the
 * reference's order, src/MinimizerSeeder.cpp:470-480). Any pointer may be NULL. */
ga_status_t ga_index_export(const ga_index_t* index, uint64_t* kmers, int64_t* starts, int64_t* counts, uint64_t* positions);

/* Seeds of n reads: MinimizerSeeder::getSeeds (src/MinimizerSeeder.cpp:521-555)
 * with the k-mer lookup on the GPU and, when order != 0,
 * GraphAligner::orderSeedsByChaining (src/GraphAligner.h:250-312): order 1 expands
 * and orders on host threads, order 2 on the GPU (one thread per read,
 * restating std::sort's tie order); both give the same seeds. *seeds is a
 * buffer from malloc (free with ga_free) holding counts[0] + ... + counts[n-1]
 * seeds, read by read, 10 int64 each in SeedHit field order
 * (src/GraphAlignerWrapper.h:11-50): nodeID, nodeOffset, seqPos, matchLen,
 * reverse, alignmentGraphNodeId, alignmentGraphNodeOffset, rawSeedGoodness,
 * seedGoodness, seedClusterSize. With order != 0, a read whose OrderSeeds fires
 * a reference assertion gets no seeds (the reference drops the read). The reads
 * are uploaded by the call, replacing the read set of ga_reads_upload. */
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
	int64_t persistentWorkers; /* deviceSeedLoop 2: workers per launch */
	double tuneMs;         /* deviceSeedLoop 2: time spent choosing the worker count (first call) */
	int64_t proposalClusters; /* learned proposal (ga_context_set_proposal): seed clusters scored, */
	int64_t prunedClusters;   /* clusters removed, */
	int64_t prunedSeeds;      /* seeds removed with them, */
	double proposalMs;        /* and the time taken (features, network, pruning) */
} ga_align_stats_t;

/* Per-read alignment of src/Aligner.cpp:390-606 (seed, order, extend in a seed
 * loop, select, GAF) for n reads with the given parameters. The seed loop of
 * each read makes the decisions of GraphAligner::AlignOneWay
 * (src/GraphAligner.h:94-181): by default the loops advance in rounds, each
 * round's extensions of all reads in one GPU launch (params->deviceSeedLoop = 1
 * runs every read's whole loop on the GPU in one launch instead). The reads are
 * uploaded by the call, replacing the read set of ga_reads_upload.
 * *gaf is a NUL-terminated buffer from malloc (free with
 * ga_free): the GAF lines (src/GraphAlignerGAFAlignment.h:37-250) of read 0,
 * then read 1, ... status[i]: 0 aligned, 1 no seeds, 2 a reference assertion
 * fired and the read was dropped, 3 no alignment kept, 4 an extension exceeded
 * the largest device workspace. stats may be NULL. */
ga_status_t ga_align_reads(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t n, const char* const* names, const char* const* seqs, const int64_t* lens, char** gaf, int32_t* status, ga_align_stats_t* stats);

/* Reads resident on the context's GPU: ga_reads_upload copies n reads into HBM
 * (and keeps a host copy for trace fix-ups and GAF text), replacing the previous
 * set; ga_align_resident then runs ga_align_reads' per-read pipeline on resident
 * reads [first, first + count) with no host-to-device read transfer. names[i]
 * names read first + i; gaf and status are as in ga_align_reads.
 * ga_align_resident fails with GA_ESTALE when a call that uploads its own reads
 * (ga_seeds, ga_align_reads, ga_align_one_way) ran after the last
 * ga_reads_upload on the context. */
ga_status_t ga_reads_upload(ga_context_t* ctx, int64_t n, const char* const* seqs, const int64_t* lens);
ga_status_t ga_align_resident(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t first, int64_t count, const char* const* names, char** gaf, int32_t* status, ga_align_stats_t* stats);

/* ---- AlignOneWay with the caller's seeds, and the result object --------- */

/* One seed, field for field SeedHit (src/GraphAlignerWrapper.h:11-50); the rows
 * ga_seeds writes have this layout. alignmentGraphNodeId and
 * alignmentGraphNodeOffset are -1 when not known (a seed from outside the
 * minimizer seeder, e.g. a seed file, src/Aligner.cpp:685-706); OrderSeeds then
 * places the seed by nodeID, reverse and nodeOffset. */
typedef struct
{
	int64_t nodeID;
	int64_t nodeOffset;
	int64_t seqPos;
	int64_t matchLen;
	int64_t reverse;
	int64_t alignmentGraphNodeId;
	int64_t alignmentGraphNodeOffset;
	int64_t rawSeedGoodness;
	int64_t seedGoodness;
	int64_t seedClusterSize;
} ga_seed_hit_t;

/* OrderSeeds (src/GraphAlignerWrapper.h:59, src/GraphAligner.h:250-312) on the
 * n seeds of one read, in place: seedGoodness and seedClusterSize are set and
 * the seeds sorted best first. GA_ASSERT when one of the reference's assertions
 * there fires (the reference drops the read); GA_EBADARG for a seed whose node
 * is not in the graph. */
ga_status_t ga_order_seeds(const ga_graph_t* graph, ga_seed_hit_t* seeds, int64_t n);

/* AlignmentResult (src/GraphAlignerCommon.h:282-344) of a batch of reads. */
typedef struct ga_alignment_result ga_alignment_result_t;

/* AlignOneWay (src/GraphAlignerWrapper.h:55, src/GraphAligner.h:94-181) for n
 * reads, each with its own ordered seeds: read i's seeds are
 * seeds[seedCounts[0] + ... + seedCounts[i-1], ... + seedCounts[i]). The
 * extensions of all reads run on the context's GPU in batched rounds; the
 * reads are uploaded by the call (replacing the read set of ga_reads_upload,
 * see GA_ESTALE). Per read the result holds AlignmentResult::alignments in the
 * order AlignOneWay leaves them (not yet selected) and seedsExtended; its
 * status is 0, 2 when a reference assertion fired (no seeds, a seed past the
 * read's end, or a check on the way; the reference drops the read) or 4 when an
 * extension exceeded the largest device workspace. params' seeding fields are
 * not used. GA_EBADARG for a seed whose nodeID/nodeOffset is not in the graph.
 * Free the result with ga_result_free. */
ga_status_t ga_align_one_way(ga_context_t* ctx, const ga_graph_t* graph, const ga_params_t* params, int64_t n, const char* const* seqs, const int64_t* lens, const ga_seed_hit_t* seeds, const int64_t* seedCounts, ga_alignment_result_t** result);

/* AlignmentResult::AlignmentItem (src/GraphAlignerCommon.h:289-340). */
typedef struct
{
	int64_t alignmentStart;
	int64_t alignmentEnd;
	int64_t alignmentScore;
	double alignmentXScore;
	int64_t seedGoodness;
	int32_t mappingQuality;
	int32_t pad;
	int64_t traceLength;
} ga_alignment_item_t;

/* TraceItem (src/GraphAlignerCommon.h:110-151) of a fixed-up trace: bigraph
 * node id (2 * GFA node + strand), offset in it, read position, node switch,
 * the read and graph letters. */
typedef struct
{
	int64_t node;
	int64_t nodeOffset;
	int64_t seqPos;
	int8_t nodeSwitch;
	char sequenceCharacter;
	char graphCharacter;
	char pad[5];
} ga_trace_item_t;

int64_t ga_result_reads(const ga_alignment_result_t* result);
/* status of read `read` (0, 2 or 4 as above), -1 for a bad argument */
int32_t ga_result_status(const ga_alignment_result_t* result, int64_t read);
int64_t ga_result_seeds_extended(const ga_alignment_result_t* result, int64_t read);
int64_t ga_result_count(const ga_alignment_result_t* result, int64_t read);
ga_status_t ga_result_item(const ga_alignment_result_t* result, int64_t read, int64_t i, ga_alignment_item_t* item);
/* Copies alignment i's trace (item.traceLength items; cap at least that many). */
ga_status_t ga_result_trace(const ga_alignment_result_t* result, int64_t read, int64_t i, ga_trace_item_t* items, int64_t cap);
/* AlignmentSelection::SelectAlignments (src/AlignmentSelection.cpp:31-95) with
 * params->minAlignmentScore and multimapScoreFraction on the read's alignments,
 * in place; with sortForOutput != 0 then the output order of
 * src/Aligner.cpp:534,549 (by start, then by X-score, best first). */
ga_status_t ga_select_alignments(ga_alignment_result_t* result, int64_t read, const ga_params_t* params, int32_t sortForOutput);
/* AlignmentSelection::AddMappingQualities (src/AlignmentSelection.cpp:97-135) on
 * the read's alignments; GA_ASSERT when one of its assertions fires. */
ga_status_t ga_add_mapping_qualities(ga_alignment_result_t* result, int64_t read);
/* AddGAFLine (src/GraphAlignerWrapper.h:57, src/GraphAlignerGAFAlignment.h:37-250):
 * the GAF line of alignment i (no newline) into buf (cap bytes with its NUL);
 * *needed is the line's length plus one. GA_EBADARG if cap is too small,
 * GA_ASSERT when a reference assertion fires. */
ga_status_t ga_gaf_line(const ga_graph_t* graph, const ga_alignment_result_t* result, int64_t read, int64_t i, const char* name, int64_t readLength, int32_t cigarMatchMismatchMerge, char* buf, int64_t cap, int64_t* needed);
void ga_result_free(ga_alignment_result_t* result);

/* ---- Learned proposal: seed-cluster network, outcome records, pruning ----
 * No counterpart in the reference: an approximate mode beside the plain method.
 * A seed cluster is a group of OrderSeeds (src/GraphAligner.h:250-312): seeds
 * of one chain whose diagonals lie within 100 of their neighbours'. The network
 * (48 features -> 128 -> 128 -> 2, ReLU, sigmoid outputs; BF16 weights and
 * activations, FP32 accumulation, gfx950 MFMA forward and backward kernels)
 * scores a cluster's chance that an alignment made from it is kept (output 0)
 * and that the plain loop extends it at all (output 1). */

#define GA_MLP_FEATURES 48
#define GA_MLP_OUTPUTS 2

typedef struct ga_mlp ga_mlp_t;
typedef struct ga_records ga_records_t;

/* Number of FP32 parameters in the flat layout of ga_mlp_get/set_params and
 * ga_mlp_gradients: W1[64][128] (input-major; inputs 48..63 padding, zero),
 * b1[128], W2[128][128], b2[128], W3[128][16] (outputs 2..15 padding), b3[16]. */
int64_t ga_mlp_num_params(void);
/* A network on the context's GPU, initialised from seed (He-uniform hidden
 * layers, Glorot-uniform output layer, zero biases). */
ga_status_t ga_mlp_create(ga_context_t* ctx, uint64_t seed, ga_mlp_t** mlp);
void ga_mlp_free(ga_mlp_t* mlp);
ga_status_t ga_mlp_get_params(ga_mlp_t* mlp, float* params);
/* Replaces the parameters (padding entries must be zero) and resets Adam. */
ga_status_t ga_mlp_set_params(ga_mlp_t* mlp, const float* params);
/* probs[n][2] = sigmoid outputs for features[n][48] (FP32, rounded to BF16). */
ga_status_t ga_mlp_forward(ga_mlp_t* mlp, int64_t n, const float* features, float* probs);
/* Mean binary cross-entropy over both outputs of n records (labels[n][2] in
 * [0, 1]) into *loss and its gradient into grads (ga_mlp_num_params floats). */
ga_status_t ga_mlp_gradients(ga_mlp_t* mlp, int64_t n, const float* features, const float* labels, double* loss, float* grads);

typedef struct
{
	int32_t epochs;   /* 1..64 */
	int32_t batch;    /* records per Adam step, 1..2^20 */
	float lr, beta1, beta2, eps;
	uint64_t seed;    /* record shuffle of every epoch */
} ga_mlp_train_params_t;

/* Adam on the GPU over n records; epochLoss[e] (params->epochs floats, may be
 * NULL) receives epoch e's mean loss, *trainMs (may be NULL) the device time. */
ga_status_t ga_mlp_train(ga_mlp_t* mlp, int64_t n, const float* features, const float* labels, const ga_mlp_train_params_t* params, float* epochLoss, double* trainMs);

/* Outcome records of the plain method on resident reads [first, first + count)
 * (ga_reads_upload; GA_ESTALE as ga_align_resident): one record per seed
 * cluster of every read whose seeds were ordered, in read order, clusters by
 * first appearance in the read's seed order. Labels: [0] an alignment extended
 * from one of the cluster's seeds was selected, [1] one was made. Any proposal
 * set on the context is not applied. */
ga_status_t ga_cluster_records(ga_context_t* ctx, const ga_graph_t* graph, const ga_index_t* index, const ga_params_t* params, int64_t first, int64_t count, ga_records_t** records);
int64_t ga_records_count(const ga_records_t* records);
const float* ga_records_features(const ga_records_t* records);  /* [count][48] */
const float* ga_records_labels(const ga_records_t* records);    /* [count][2] */
const int64_t* ga_records_reads(const ga_records_t* records);   /* [count], read index - first */
void ga_records_free(ga_records_t* records);

/* From now on ga_align_reads and ga_align_resident on ctx remove, before each
 * read's seed loop, the seeds of every cluster whose output 0 is below
 * threshold, except the read's first keepTop (>= 1) clusters. Only a read's
 * first maxClusters clusters are scored; the ones after them are removed
 * unscored (maxClusters 0: every cluster is scored). mlp NULL turns the
 * proposal off (the plain method). The network must stay alive while set. */
ga_status_t ga_context_set_proposal(ga_context_t* ctx, ga_mlp_t* mlp, double threshold, int32_t keepTop, int32_t maxClusters);

void ga_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
