// Learned proposal of the seed loop: features of every seed cluster of a read,
// outcome records of the plain method, and pruning of the clusters a scorer
// (the BF16 MLP) rates below a threshold.
//
// A seed cluster is a group of OrderSeeds (src/GraphAligner.h:250-312): seeds of
// one chain whose diagonals lie within 100 of their neighbours'. The plain
// method extends a cluster's seeds in goodness order until an alignment covers
// them; the proposal removes whole clusters from a read's ordered seeds before
// its loop starts, so the loop is the plain one on fewer seeds. Results then
// differ from the plain method's only where a removed cluster would have made
// an alignment that survives selection, or would have covered (and so skipped)
// another cluster's seeds.
#pragma once
#include <cstdint>
#include <mutex>
#include <vector>
#include "ga_align.hpp"

namespace ga {

constexpr int kClusterFeatures = 48;
// 0: an alignment made from one of the cluster's seeds is in the read's
// selected output; 1: the plain loop made an alignment from one of its seeds
constexpr int kProposalOutputs = 2;

// Cluster of every seed (seeds in loop order), numbered by first appearance;
// returns the number of clusters.
uint32_t seedClusters(const GraphHost& g, const std::vector<SeedHit>& seeds, std::vector<uint32_t>& clusterOf);

// kClusterFeatures floats per cluster (nclusters rows) of one read: hit count,
// covered bases, goodness and rank, diagonal spread, read span, k-mer
// repetitiveness, split-node in/out degrees and component sizes, local k-mer
// entropy and homopolymer runs, overlap with better clusters. maxCount is the
// index's count cut-off (rawSeedGoodness = maxCount - k-mer count). Rows are
// written for clusters [0, rows) (the read's totals still count every cluster).
void clusterFeatures(const GraphHost& g, uint64_t maxCount, const char* seq, uint32_t len, const std::vector<SeedHit>& seeds, const std::vector<uint32_t>& clusterOf, uint32_t nclusters, uint32_t rows, float* out);

// Scores n feature rows: kProposalOutputs probabilities per row.
struct ClusterScorer
{
	virtual ~ClusterScorer() = default;
	virtual void score(const float* features, size_t n, float* out) = 0;
};

struct SeedProposal
{
	ClusterScorer* scorer = nullptr;
	// clusters with output 0 below threshold are removed, except the read's
	// first keepTop clusters; only the first maxClusters are scored, the ones
	// after them are removed unscored (0: every cluster is scored)
	double threshold = 0.5;
	uint32_t keepTop = 1;
	uint32_t maxClusters = 0;
};

struct ProposalStats
{
	// clusters scored, clusters removed (scored low or past maxClusters), their seeds
	uint64_t clusters = 0, prunedClusters = 0, prunedSeeds = 0;
	double ms = 0;
};

// Removes the pruned clusters' seeds from seeds[i] (order kept) for every read
// i; one scorer call for all reads.
void applyProposal(const SeedProposal& p, const GraphHost& g, uint64_t maxCount, size_t nreads, std::vector<SeedHit>* const* seeds, const char* const* seqs, const uint32_t* lens, int threads, ProposalStats& stats);

// Outcome records of the plain method: per cluster, its features and labels.
struct ClusterRecords
{
	std::mutex m;
	std::vector<float> features;    // kClusterFeatures per record
	std::vector<float> labels;      // kProposalOutputs per record
	std::vector<int64_t> reads;     // read index (first read of the call = 0)
	std::vector<int32_t> clusters;  // cluster number within the read
	// labels of one read's clusters from the alignments its loop made (all) and
	// the selected ones (kept); each identifies its seed by seedIndex
	void add(int64_t read, const std::vector<float>& feats, uint32_t nclusters, const std::vector<uint32_t>& clusterOf, const std::vector<Alignment>& all, const std::vector<Alignment>& kept);
	// records in read order, clusters in order within a read
	void sort();
};

} // namespace ga
