// Seed expansion and OrderSeeds of one read on the device (seedOrderKernel) and,
// from the same source, on the host for the CPU suite: MinimizerSeeder::getSeeds'
// expansion of a read's matches (src/MinimizerSeeder.cpp:521-555) and
// GraphAligner::orderSeedsByChaining (src/GraphAligner.h:250-312).
//
// Two of the reference's steps use std::sort, which is not stable, on keys
// that tie often (matches by count, seeds by goodness): the order of tied
// elements is whatever libstdc++'s introsort leaves, and seed order decides which
// seeds a read extends. libstdSort below restates that algorithm's steps so the
// device gives the same permutation: introsort with a depth limit of twice the
// floor of log2(n), median-of-three pivot moved to the front, unguarded
// partition, heap sort when the depth limit runs out, ranges of at most 16
// elements left to a final insertion sort. The chain and cluster steps in
// between do not depend on the order of ties (a cluster's covered bases and size
// are the same for any order of its seeds), so they use their own order.
#pragma once
#include <stdint.h>
#include "ga_device_types.h"

#if defined(__HIPCC__)
#define GA_ORDER_HD __host__ __device__ inline
#else
#define GA_ORDER_HD inline
#endif

namespace gad {

// SeedHit (src/GraphAlignerWrapper.h:11-50) as the device writes it
struct DevSeedHit
{
	uint64_t rawSeedGoodness;
	uint64_t seedGoodness;
	uint64_t seedClusterSize;
	int32_t nodeID;      // bigraph node id / 2
	uint32_t nodeOffset; // offset in the bigraph node
	uint32_t seqPos;
	uint32_t matchLen;
	uint32_t alignmentGraphNodeId;
	uint32_t alignmentGraphNodeOffset;
	uint32_t reverse;
	uint32_t pad;
};

// graph arrays the seeding reads
struct SeedGraphView
{
	const int32_t* nodeId;
	const uint32_t* nodeOff;
	const uint8_t* reverse;
	const uint32_t* chain;
	const uint64_t* chainPos;
};

template <typename T>
GA_ORDER_HD void orderSwap(T& a, T& b)
{
	T t = a;
	a = b;
	b = t;
}

template <typename T, typename Less>
GA_ORDER_HD void heapAdjust(T* first, int64_t hole, int64_t len, T value, Less less)
{
	const int64_t top = hole;
	int64_t child = hole;
	while (child < (len - 1) / 2)
	{
		child = 2 * (child + 1);
		if (less(first[child], first[child - 1])) child--;
		first[hole] = first[child];
		hole = child;
	}
	if ((len & 1) == 0 && child == (len - 2) / 2)
	{
		child = 2 * (child + 1);
		first[hole] = first[child - 1];
		hole = child - 1;
	}
	// push the value back up towards top
	int64_t parent = (hole - 1) / 2;
	while (hole > top && less(first[parent], value))
	{
		first[hole] = first[parent];
		hole = parent;
		parent = (hole - 1) / 2;
	}
	first[hole] = value;
}

// heap sort of [first, first + len): make a max-heap, then pop it from the back
template <typename T, typename Less>
GA_ORDER_HD void heapSortRange(T* first, int64_t len, Less less)
{
	if (len < 2) return;
	for (int64_t parent = (len - 2) / 2;; parent--)
	{
		heapAdjust(first, parent, len, first[parent], less);
		if (parent == 0) break;
	}
	for (int64_t last = len - 1; last > 0; last--)
	{
		T value = first[last];
		first[last] = first[0];
		heapAdjust(first, (int64_t)0, last, value, less);
	}
}

// insertion of a[i] into the sorted a[0, i) when a[0] is known not to exceed it
template <typename T, typename Less>
GA_ORDER_HD void unguardedInsert(T* a, int64_t i, Less less)
{
	T value = a[i];
	int64_t j = i - 1;
	while (less(value, a[j]))
	{
		a[i] = a[j];
		i = j;
		j--;
	}
	a[i] = value;
}

template <typename T, typename Less>
GA_ORDER_HD void insertionSortRange(T* a, int64_t n, Less less)
{
	for (int64_t i = 1; i < n; i++)
	{
		if (less(a[i], a[0]))
		{
			T value = a[i];
			for (int64_t j = i; j > 0; j--) a[j] = a[j - 1];
			a[0] = value;
		}
		else unguardedInsert(a, i, less);
	}
}

// a range of libstdSort waiting to be partitioned, with its remaining depth
struct SortRange
{
	int64_t first, last;
	int32_t depth;
};
// libstdSort keeps at most floor(log2(n)) ranges pending, below 64 for any n
constexpr int kSortStackEntries = 64;

// std::sort's permutation of a[0, n) under less (see the header comment). The
// recursion becomes a loop over the smaller part of each partition with the
// larger part pushed on stack (kSortStackEntries entries, so at most log2(n) are
// pending): ranges are disjoint, so the order in which they are partitioned does
// not change the result.
template <typename T, typename Less>
GA_ORDER_HD void libstdSort(T* a, int64_t n, Less less, SortRange* stack)
{
	if (n < 2) return;
	const int64_t threshold = 16;
	int lg = 0;
	while (((uint64_t)n >> (lg + 1)) != 0) lg++;
	int top = 0;
	stack[top++] = SortRange{0, n, 2 * lg};
	while (top > 0)
	{
		SortRange r = stack[--top];
		int64_t first = r.first, last = r.last;
		int32_t depth = r.depth;
		while (last - first > threshold)
		{
			if (depth == 0)
			{
				heapSortRange(a + first, last - first, less);
				break;
			}
			depth--;
			// median of first + 1, middle and last - 1 moved to first
			const int64_t mid = first + (last - first) / 2;
			const int64_t x = first + 1, y = mid, z = last - 1;
			int64_t pick;
			if (less(a[x], a[y]))
			{
				if (less(a[y], a[z])) pick = y;
				else if (less(a[x], a[z])) pick = z;
				else pick = x;
			}
			else if (less(a[x], a[z])) pick = x;
			else if (less(a[y], a[z])) pick = z;
			else pick = y;
			orderSwap(a[first], a[pick]);
			// unguarded partition of [first + 1, last) around a[first]
			int64_t lo = first + 1, hi = last;
			while (true)
			{
				while (less(a[lo], a[first])) lo++;
				hi--;
				while (less(a[first], a[hi])) hi--;
				if (!(lo < hi)) break;
				orderSwap(a[lo], a[hi]);
				lo++;
			}
			// [first, lo) and [lo, last) both go on with this depth
			if (lo - first < last - lo)
			{
				stack[top++] = SortRange{lo, last, depth};
				last = lo;
			}
			else
			{
				stack[top++] = SortRange{first, lo, depth};
				first = lo;
			}
		}
	}
	// final insertion sort: guarded over the first 16, unguarded after them
	if (n > threshold)
	{
		insertionSortRange(a, threshold, less);
		for (int64_t i = threshold; i < n; i++) unguardedInsert(a, i, less);
	}
	else insertionSortRange(a, n, less);
}

// (chain, diagonal) of one seed, for the chain and cluster steps
struct ChainKey
{
	uint64_t chain;
	uint64_t diagonal;
	uint32_t seed;
	uint32_t seqPos;
};

// Seeds of one read from its matches (sorted in place by count), ordered as
// OrderSeeds leaves them, into out[0, cap). Returns the number of seeds, -1 when
// a reference assertion of OrderSeeds fires (the read is dropped), -2 when cap
// is too small, -3 when a match lies outside the npositions positions. scratch
// holds cap entries, stack kSortStackEntries.
GA_ORDER_HD int64_t seedsOfRead(const SeedGraphView& g, const uint64_t* positions, uint64_t npositions, uint64_t maxCount, uint32_t k, SeedMatch* m,
                                uint32_t nm, uint64_t readLen, double density, DevSeedHit* out, uint64_t cap, ChainKey* scratch, SortRange* stack)
{
	for (uint32_t i = 0; i < nm; i++)
		if ((uint64_t)m[i].start + m[i].count > npositions) return -3;
	libstdSort(m, (int64_t)nm, [](const SeedMatch& l, const SeedMatch& r) { return l.count < r.count; }, stack);
	const uint64_t maxHits = density == -1 ? ~(uint64_t)0 : (uint64_t)(readLen * density);
	uint64_t n = 0, seedsHere = 0, allowed = 0;
	for (uint32_t i = 0; i < nm; i++)
	{
		const uint64_t count = m[i].count;
		if (seedsHere >= maxHits && count > allowed) break;
		allowed = count;
		if (n + count > cap) return -2;
		for (uint32_t c = 0; c < m[i].count; c++)
		{
			const uint64_t merged = positions[(uint64_t)m[i].start + c];
			const uint32_t split = (uint32_t)(merged >> 6);
			const uint32_t offset = (uint32_t)(merged & 63);
			DevSeedHit h;
			h.nodeID = g.nodeId[split] / 2;
			h.nodeOffset = offset + g.nodeOff[split];
			h.seqPos = m[i].pos;
			h.matchLen = k;
			h.rawSeedGoodness = maxCount - count;
			h.seedGoodness = 0;
			h.seedClusterSize = 0;
			h.reverse = g.reverse[split] != 0;
			h.alignmentGraphNodeId = split;
			h.alignmentGraphNodeOffset = offset;
			h.pad = 0;
			out[n++] = h;
		}
		seedsHere += count;
	}
	// chains and clusters (diagonals within 100 of the previous one)
	for (uint64_t i = 0; i < n; i++)
	{
		const uint32_t node = out[i].alignmentGraphNodeId;
		const uint64_t d = g.chainPos[node] + out[i].alignmentGraphNodeOffset;
		if (d < out[i].seqPos) return -1;
		scratch[i] = ChainKey{g.chain[node], d - out[i].seqPos, (uint32_t)i, out[i].seqPos};
	}
	libstdSort(scratch, (int64_t)n, [](const ChainKey& l, const ChainKey& r) {
		return l.chain != r.chain ? l.chain < r.chain : l.diagonal != r.diagonal ? l.diagonal < r.diagonal : l.seed < r.seed;
	}, stack);
	uint64_t start = 0;
	for (uint64_t i = 1; i <= n; i++)
	{
		if (i < n && scratch[i].chain == scratch[i - 1].chain && scratch[i].diagonal <= scratch[i - 1].diagonal + 100) continue;
		// cluster [start, i): bases covered by its seeds' k-mers, by read position
		libstdSort(scratch + start, (int64_t)(i - start), [](const ChainKey& l, const ChainKey& r) { return l.seqPos != r.seqPos ? l.seqPos < r.seqPos : l.seed < r.seed; }, stack);
		uint64_t covered = 0;
		int32_t lastEnd = INT32_MIN;
		for (uint64_t c = start; c < i; c++)
		{
			const DevSeedHit& s = out[scratch[c].seed];
			const int32_t thisStart = (int32_t)s.seqPos - (int32_t)s.matchLen + 1;
			const int32_t thisEnd = (int32_t)s.seqPos;
			if (thisEnd < lastEnd || thisEnd <= thisStart) return -1;
			covered += (uint64_t)(thisEnd - (thisStart > lastEnd ? thisStart : lastEnd));
			lastEnd = thisEnd;
		}
		for (uint64_t c = start; c < i; c++)
		{
			DevSeedHit& s = out[scratch[c].seed];
			s.seedGoodness = covered + s.rawSeedGoodness;
			s.seedClusterSize = i - start;
		}
		start = i;
	}
	libstdSort(out, (int64_t)n, [](const DevSeedHit& l, const DevSeedHit& r) { return l.seedGoodness < r.seedGoodness; }, stack);
	for (uint64_t i = 0, j = n; i + 1 < j; i++, j--) orderSwap(out[i], out[j - 1]);
	return (int64_t)n;
}

} // namespace gad
