// Minimizer index over the alignment graph (MinimizerSeeder, src/MinimizerSeeder.cpp)
// laid out for HBM: an open-addressing table k-mer -> (start, count) into a position
// array, positions (splitNode << 6 | offset) of one k-mer stored in the order the
// reference's single-bucket build leaves them (reverse discovery order).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "ga_graph.hpp"

namespace ga {

struct KmerSlot
{
	uint64_t kmer;  // ~0 = empty
	uint32_t start;
	uint32_t count;
};

struct MinimizerIndex
{
	uint64_t uid = 0;
	uint32_t k = 15, w = 20;
	uint64_t maxCount = 0;
	uint64_t mask = 0;                // table size - 1 (power of two)
	std::vector<KmerSlot> table;
	std::vector<uint64_t> positions;
	uint64_t numKmers = 0;
	const KmerSlot* find(uint64_t kmer) const;
};

// src/MinimizerSeeder.cpp:299-491,557-575 (see ga_index.cpp for the maxCount rule)
MinimizerIndex buildMinimizerIndex(const GraphHost& g, uint32_t k, uint32_t w, double discardFraction);

class Context;
// The same index built on the context's GPU (ga_index_gpu.hip): identical
// k-mers, ranges, position order and maxCount; table slots may differ.
MinimizerIndex buildMinimizerIndexGpu(Context& ctx, const GraphHost& g, uint32_t k, uint32_t w, double discardFraction);

} // namespace ga
