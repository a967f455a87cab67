// Minimizer index built on the GPU: the same index as buildMinimizerIndex
// (ga_index.cpp; MinimizerSeeder's constructor, src/MinimizerSeeder.cpp:299-491,
// and initMaxCount, :557-575), computed as data-parallel passes over the bases of
// every bigraph node instead of a sequential window walk per node:
//
//  1. every split node writes its 2-bit codes into one buffer holding the bigraph
//     nodes in the lookup map's order (one invalid code between nodes), and each
//     bigraph node's minimizer start (DBG overlaps) is the largest offset of a
//     split piece with an in-edge from another bigraph node;
//  2. every position hashes the k-mer ending there;
//  3. every position decides from the window minima around it whether the
//     reference's window walk (iterateMinimizersReal, :104-189) reports it: a
//     run of valid bases reports nothing until it is w + 1 long; its first window
//     holds the r + 1 k-mers ending at s0 + k - 1 .. s0 + w (r = w - k + 1) and
//     reports every minimum position; each later window ending at e (r k-mers)
//     reports e if e holds its minimum, and, when its minimum rose (the old one
//     left), the last earlier position holding the new minimum. The walk reports
//     positions in increasing order, so buffer order is its discovery order;
//  4. the reported (k-mer, position) pairs are compacted in that order, sorted by
//     (k-mer, discovery index), grouped, and laid out as the host build does:
//     groups in order of first discovery, each group's positions in reverse
//     discovery order, and an open-addressing table k-mer -> (start, count);
//  5. maxCount from a histogram of the group sizes.
//
// Table slots are claimed with 64-bit compare-and-swap, so a k-mer may sit in
// another slot than in the host build; every lookup returns the same range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "ga_index.hpp"
#include "ga_runtime.hpp"
#include "ga_seed_core.h"

namespace gad {

constexpr uint8_t kNoBase = 4;
constexpr uint32_t kScanBlock = 1024; // items per scan workgroup (256 threads x 4)

__global__ void __launch_bounds__(256) idxFillKernel(const uint32_t* __restrict__ nodeLen, const uint64_t* __restrict__ seq2, const uint64_t* __restrict__ amb4,
                                                     uint32_t firstAmbiguous, const uint64_t* __restrict__ nodeBufOff, uint32_t numNodes,
                                                     uint8_t* __restrict__ code, uint32_t* __restrict__ splitOf)
{
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= numNodes) return;
	const uint64_t off = nodeBufOff[n];
	const uint32_t L = nodeLen[n];
	for (uint32_t p = 0; p < L; p++)
	{
		uint8_t c;
		if (n < firstAmbiguous) c = (uint8_t)((seq2[2 * (uint64_t)n + (p >> 5)] >> ((p & 31) * 2)) & 3);
		else
		{
			// an IUPAC node's letter is a base only when exactly one mask has it
			const uint64_t* m = amb4 + 4 * (uint64_t)(n - firstAmbiguous);
			const uint32_t set = (uint32_t)((m[0] >> p) & 1) | (uint32_t)(((m[1] >> p) & 1) << 1) | (uint32_t)(((m[2] >> p) & 1) << 2) | (uint32_t)(((m[3] >> p) & 1) << 3);
			c = set == 1 ? 0 : set == 2 ? 1 : set == 4 ? 2 : set == 8 ? 3 : kNoBase;
		}
		code[off + p] = c;
		splitOf[off + p] = n;
	}
}

// minimizer start of each bigraph node (src/MinimizerSeeder.cpp:320-337)
__global__ void __launch_bounds__(256) idxMinStartKernel(const uint32_t* __restrict__ inOff, const uint32_t* __restrict__ inAdj, const int32_t* __restrict__ nodeId,
                                                         const uint32_t* __restrict__ nodeOff, const uint32_t* __restrict__ nodeRank, uint32_t numNodes,
                                                         uint32_t* __restrict__ minStart)
{
	const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= numNodes) return;
	for (uint32_t e = inOff[n]; e < inOff[n + 1]; e++)
	{
		if (nodeId[inAdj[e]] != nodeId[n])
		{
			atomicMax(&minStart[nodeRank[n]], nodeOff[n]);
			break;
		}
	}
}

__global__ void __launch_bounds__(256) idxHashKernel(const uint8_t* __restrict__ code, uint64_t nbases, uint32_t k, uint64_t* __restrict__ kmerOut,
                                                     uint64_t* __restrict__ hashOut, uint8_t* __restrict__ valid)
{
	const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nbases) return;
	uint8_t ok = e + 1 >= k;
	uint64_t kmer = 0;
	for (uint32_t i = 0; ok && i < k; i++)
	{
		const uint8_t c = code[e + 1 - k + i];
		if (c >= kNoBase) ok = 0;
		kmer = (kmer << 2) | c;
	}
	valid[e] = ok;
	kmerOut[e] = ok ? kmer : 0;
	hashOut[e] = ok ? seedKmerHash(kmer) : 0;
}

__device__ __forceinline__ uint64_t minHash(const uint64_t* h, uint64_t lo, uint64_t hi)
{
	uint64_t m = ~(uint64_t)0;
	for (uint64_t i = lo; i <= hi; i++) m = h[i] < m ? h[i] : m;
	return m;
}

// Step 3 of the header: ownFlag[e] for the reports made at e itself, laterFlag[j]
// = 1 for a position a later window reports (plain stores of 1 only).
__global__ void __launch_bounds__(256) idxWindowKernel(const uint8_t* __restrict__ code, const uint8_t* __restrict__ valid, const uint64_t* __restrict__ h, uint64_t nbases,
                                                       uint32_t k, uint32_t w, uint8_t* __restrict__ ownFlag, uint8_t* __restrict__ laterFlag)
{
	const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nbases) return;
	ownFlag[e] = 0;
	if (!valid[e]) return;
	const uint64_t r = w - k + 1;
	// start s0 of the run of valid bases holding e (the walk restarts after an
	// invalid base): needed only when e <= s0 + w + 1, i.e. when an invalid base
	// (or the buffer start) lies at most w + 2 bases before e
	const uint64_t back = e >= (uint64_t)w + 2 ? e - (w + 2) : 0;
	int64_t s0 = -1;
	for (uint64_t i = e + 1; i-- > back;)
	{
		if (code[i] >= kNoBase)
		{
			s0 = (int64_t)i + 1;
			break;
		}
	}
	if (s0 < 0 && e < (uint64_t)w + 2) s0 = 0;
	if (s0 >= 0 && e <= (uint64_t)s0 + w)
	{
		// first window: the run must reach s0 + w
		const uint64_t end = (uint64_t)s0 + w;
		if (end >= nbases) return;
		for (uint64_t i = e + 1; i <= end; i++)
			if (code[i] >= kNoBase) return;
		ownFlag[e] = h[e] == minHash(h, (uint64_t)s0 + k - 1, end);
		return;
	}
	// a later window ending at e (its r k-mers are valid: the run started at least
	// w + 1 bases before e)
	const uint64_t m = minHash(h, e - r + 1, e);
	ownFlag[e] = h[e] == m;
	const uint64_t prev = (s0 >= 0 && e == (uint64_t)s0 + w + 1) ? minHash(h, (uint64_t)s0 + k - 1, (uint64_t)s0 + w) : minHash(h, e - r, e - 1);
	if (m > prev)
	{
		for (uint64_t j = e - 1; j + r > e; j--)
		{
			if (h[j] == m)
			{
				laterFlag[j] = 1;
				break;
			}
		}
	}
}

__global__ void __launch_bounds__(256) idxKeepKernel(const uint8_t* __restrict__ ownFlag, const uint8_t* __restrict__ laterFlag, const uint32_t* __restrict__ splitOf,
                                                     const uint64_t* __restrict__ nodeBufOff, const uint32_t* __restrict__ nodeOff,
                                                     const uint32_t* __restrict__ nodeRank, const uint32_t* __restrict__ minStart, uint64_t nbases,
                                                     uint32_t* __restrict__ keep)
{
	const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nbases) return;
	uint32_t f = ownFlag[e] | laterFlag[e];
	if (f)
	{
		const uint32_t s = splitOf[e];
		const uint64_t posInBig = nodeOff[s] + (e - nodeBufOff[s]);
		if (posInBig < minStart[nodeRank[s]]) f = 0;
	}
	keep[e] = f;
}

// exclusive scan of 1024 items per workgroup; blockSums[b] = the block's total
__global__ void __launch_bounds__(256) scanBlockKernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n, uint32_t* __restrict__ blockSums)
{
	__shared__ uint32_t waveTot[4];
	const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + 4 * threadIdx.x;
	uint32_t v[4];
	uint32_t sum = 0;
	for (int i = 0; i < 4; i++)
	{
		v[i] = base + i < n ? in[base + i] : 0;
		sum += v[i];
	}
	// inclusive scan of the threads' sums within the wave
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint32_t inc = sum;
	for (int d = 1; d < 64; d <<= 1)
	{
		const uint32_t t = __shfl_up(inc, d, 64);
		if (lane >= d) inc += t;
	}
	if (lane == 63) waveTot[wave] = inc;
	__syncthreads();
	uint32_t waveOff = 0;
	for (int i = 0; i < wave; i++) waveOff += waveTot[i];
	uint32_t run = waveOff + inc - sum;
	for (int i = 0; i < 4; i++)
	{
		if (base + i < n) out[base + i] = run;
		run += v[i];
	}
	if (threadIdx.x == 255) blockSums[blockIdx.x] = waveOff + inc;
}

__global__ void __launch_bounds__(256) scanAddKernel(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ blockOff)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] += blockOff[i / kScanBlock];
}

__global__ void __launch_bounds__(256) idxCompactKernel(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ slot, const uint64_t* __restrict__ kmer,
                                                        const uint32_t* __restrict__ splitOf, const uint64_t* __restrict__ nodeBufOff, uint64_t nbases,
                                                        uint64_t* __restrict__ keyOut, uint32_t* __restrict__ idxOut, uint64_t* __restrict__ posOut)
{
	const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nbases || !keep[e]) return;
	const uint32_t i = slot[e];
	const uint32_t s = splitOf[e];
	keyOut[i] = kmer[e];
	idxOut[i] = i;
	posOut[i] = ((uint64_t)s << 6) + (e - nodeBufOff[s]);
}

__global__ void __launch_bounds__(256) padKeysKernel(uint64_t* __restrict__ key, uint32_t* __restrict__ val, uint32_t from, uint32_t to)
{
	const uint32_t i = from + blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= to) return;
	key[i] = ~(uint64_t)0;
	val[i] = ~0u;
}

// one compare-exchange step of a bitonic sort of (key, val) pairs, ascending
__global__ void __launch_bounds__(256) bitonicStepKernel(uint64_t* __restrict__ key, uint32_t* __restrict__ val, uint32_t n, uint32_t size, uint32_t stride)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t j = i ^ stride;
	if (j <= i) return;
	const uint64_t ki = key[i], kj = key[j];
	const uint32_t vi = val[i], vj = val[j];
	const bool greater = ki > kj || (ki == kj && vi > vj);
	const bool ascending = (i & size) == 0;
	if (greater == ascending)
	{
		key[i] = kj;
		key[j] = ki;
		val[i] = vj;
		val[j] = vi;
	}
}

__global__ void __launch_bounds__(256) idxHeadsKernel(const uint64_t* __restrict__ key, uint32_t n, uint32_t* __restrict__ head)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// per group: where it starts in sorted order; and a mark at its first discovery index
__global__ void __launch_bounds__(256) idxGroupsKernel(const uint32_t* __restrict__ head, const uint32_t* __restrict__ headScan, const uint32_t* __restrict__ val, uint32_t n,
                                                       uint32_t* __restrict__ groupStart, uint32_t* __restrict__ firstMark)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || !head[i]) return;
	groupStart[headScan[i]] = i;
	firstMark[val[i]] = 1;
}

__global__ void __launch_bounds__(256) idxGroupRankKernel(const uint32_t* __restrict__ groupStart, uint32_t groups, uint32_t n, const uint32_t* __restrict__ val,
                                                          const uint32_t* __restrict__ firstScan, uint32_t* __restrict__ groupRank,
                                                          uint32_t* __restrict__ countByRank, uint32_t* __restrict__ maxCount)
{
	const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= groups) return;
	const uint32_t s = groupStart[g];
	const uint32_t c = (g + 1 < groups ? groupStart[g + 1] : n) - s;
	const uint32_t rank = firstScan[val[s]];
	groupRank[g] = rank;
	countByRank[rank] = c;
	atomicMax(maxCount, c);
}

__global__ void __launch_bounds__(256) idxPlaceKernel(const uint32_t* __restrict__ head, const uint32_t* __restrict__ val, const uint32_t* __restrict__ headScan, uint32_t n,
                                                      const uint32_t* __restrict__ groupStart, uint32_t groups, const uint32_t* __restrict__ groupRank,
                                                      const uint32_t* __restrict__ startByRank, const uint64_t* __restrict__ posFound,
                                                      uint64_t* __restrict__ positions)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	// group of sorted element i: the heads up to and including i, less one
	const uint32_t g = headScan[i] + head[i] - 1;
	const uint32_t s = groupStart[g];
	const uint32_t c = (g + 1 < groups ? groupStart[g + 1] : n) - s;
	const uint32_t start = startByRank[groupRank[g]];
	// reverse discovery order within the k-mer
	positions[start + c - 1 - (i - s)] = posFound[val[i]];
}

// group g's slot in the table (linear probing, slots claimed by compare-and-swap)
// and its count in the histogram
__device__ __forceinline__ void idxInsertGroup(const uint64_t* __restrict__ key, const uint32_t* __restrict__ groupStart, uint32_t groups, uint32_t n,
                                               const uint32_t* __restrict__ groupRank, const uint32_t* __restrict__ startByRank, SeedTableSlot* table,
                                               uint64_t mask, uint32_t* __restrict__ histogram, uint32_t histSize, uint32_t* local, uint32_t g)
{
	const uint32_t s = groupStart[g];
	const uint32_t c = (g + 1 < groups ? groupStart[g + 1] : n) - s;
	const uint64_t kmer = key[s];
	// the table holds at least twice as many slots as k-mers: a free slot is found
	uint64_t i = seedKmerHash(kmer) & mask;
	while (true)
	{
		const unsigned long long old = atomicCAS((unsigned long long*)&table[i].kmer, ~0ull, (unsigned long long)kmer);
		if (old == ~0ull) break;
		i = (i + 1) & mask;
	}
	table[i].start = startByRank[groupRank[g]];
	table[i].count = c;
	if (c < 256) atomicAdd(&local[c], 1u);
	else if (c < histSize) atomicAdd(&histogram[c], 1u);
}

__global__ void __launch_bounds__(256) idxTableKernel(const uint64_t* __restrict__ key, const uint32_t* __restrict__ groupStart, uint32_t groups, uint32_t n,
                                                      const uint32_t* __restrict__ groupRank, const uint32_t* __restrict__ startByRank, SeedTableSlot* table,
                                                      uint64_t mask, uint32_t* __restrict__ histogram, uint32_t histSize)
{
	// counts below 256 (nearly all k-mers) are tallied in LDS first: one global
	// atomic per bin and workgroup instead of one per k-mer on a few hot bins
	__shared__ uint32_t local[256];
	local[threadIdx.x] = 0;
	__syncthreads();
	const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < groups) idxInsertGroup(key, groupStart, groups, n, groupRank, startByRank, table, mask, histogram, histSize, local, g);
	__syncthreads();
	if (local[threadIdx.x] != 0 && threadIdx.x < histSize) atomicAdd(&histogram[threadIdx.x], local[threadIdx.x]);
}

} // namespace gad

namespace ga {

namespace {

uint32_t blocksFor(uint64_t n, uint32_t per = 256) { return (uint32_t)((n + per - 1) / per); }

// exclusive scan of n u32 values (in may equal out); returns the total
uint32_t scanU32(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, std::vector<std::unique_ptr<DeviceBuffer>>& keep)
{
	if (n == 0) return 0;
	const uint64_t nb = (n + gad::kScanBlock - 1) / gad::kScanBlock;
	keep.push_back(std::make_unique<DeviceBuffer>());
	DeviceBuffer& sums = *keep.back();
	sums.reserve((nb + 1) * sizeof(uint32_t));
	hipLaunchKernelGGL(gad::scanBlockKernel, dim3((uint32_t)nb), dim3(256), 0, st, in, out, n, sums.as<uint32_t>());
	GA_HIP(hipGetLastError());
	uint32_t total = 0;
	if (nb > 1)
	{
		total = scanU32(st, sums.as<uint32_t>(), sums.as<uint32_t>(), nb, keep);
		hipLaunchKernelGGL(gad::scanAddKernel, dim3(blocksFor(n)), dim3(256), 0, st, out, n, sums.as<uint32_t>());
		GA_HIP(hipGetLastError());
	}
	else
	{
		GA_HIP(hipMemcpyAsync(&total, sums.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		GA_HIP(hipStreamSynchronize(st));
	}
	return total;
}

template <typename T>
T* devAlloc(std::vector<std::unique_ptr<DeviceBuffer>>& keep, uint64_t count)
{
	keep.push_back(std::make_unique<DeviceBuffer>());
	keep.back()->reserve(std::max<uint64_t>(count, 1) * sizeof(T));
	return keep.back()->as<T>();
}

} // namespace

MinimizerIndex buildMinimizerIndexGpu(Context& ctx, const GraphHost& g, uint32_t k, uint32_t w, double discardFraction)
{
	if (k < 1 || k > 31 || w < k) throw std::runtime_error("minimizer length must be 1..31 and not above the window size");
	GA_HIP(hipSetDevice(ctx.device()));
	hipStream_t st = ctx.stream();
	const gad::DevGraph& dg = ctx.deviceGraph(g);
	const uint32_t N = (uint32_t)g.numNodes();
	// buffer layout: bigraph nodes in the lookup map's order, one invalid code apart
	std::unordered_map<int, uint32_t> rankOf;
	std::vector<uint64_t> idBase(g.lookupOrder.size());
	{
		rankOf.reserve(g.lookupOrder.size());
		uint64_t at = 0;
		for (size_t r = 0; r < g.lookupOrder.size(); r++)
		{
			rankOf[g.lookupOrder[r]] = (uint32_t)r;
			idBase[r] = at;
			at += g.originalSize.at(g.lookupOrder[r]) + 1;
		}
		if (at >= (1ull << 32)) throw std::runtime_error("graph too large for the GPU index build (4 Gbp)");
	}
	const uint64_t nbases = g.lookupOrder.empty() ? 0 : idBase.back() + g.originalSize.at(g.lookupOrder.back()) + 1;
	std::vector<uint64_t> nodeBufOff(N);
	std::vector<uint32_t> nodeRank(N);
	for (uint32_t n = 0; n < N; n++)
	{
		nodeRank[n] = rankOf.at(g.nodeId[n]);
		nodeBufOff[n] = idBase[nodeRank[n]] + g.nodeOff[n];
	}
	std::vector<std::unique_ptr<DeviceBuffer>> keep;
	uint64_t* dBufOff = devAlloc<uint64_t>(keep, N);
	uint32_t* dRank = devAlloc<uint32_t>(keep, N);
	uint32_t* dMinStart = devAlloc<uint32_t>(keep, g.lookupOrder.size());
	uint8_t* dCode = devAlloc<uint8_t>(keep, nbases);
	uint32_t* dSplit = devAlloc<uint32_t>(keep, nbases);
	GA_HIP(hipMemcpyAsync(dBufOff, nodeBufOff.data(), N * sizeof(uint64_t), hipMemcpyHostToDevice, st));
	GA_HIP(hipMemcpyAsync(dRank, nodeRank.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, st));
	GA_HIP(hipMemsetAsync(dMinStart, 0, std::max<size_t>(g.lookupOrder.size(), 1) * sizeof(uint32_t), st));
	GA_HIP(hipMemsetAsync(dCode, gad::kNoBase, std::max<uint64_t>(nbases, 1), st));
	MinimizerIndex idx;
	idx.uid = nextObjectUid();
	idx.k = k;
	idx.w = w;
	if (N > 0)
	{
		hipLaunchKernelGGL(gad::idxFillKernel, dim3(blocksFor(N)), dim3(256), 0, st, dg.nodeLen, dg.seq2, dg.amb4, dg.firstAmbiguous, dBufOff, N, dCode, dSplit);
		hipLaunchKernelGGL(gad::idxMinStartKernel, dim3(blocksFor(N)), dim3(256), 0, st, dg.inOff, dg.inAdj, dg.nodeId, dg.nodeOff, dRank, N, dMinStart);
		GA_HIP(hipGetLastError());
	}
	uint64_t* dKmer = devAlloc<uint64_t>(keep, nbases);
	uint64_t* dHash = devAlloc<uint64_t>(keep, nbases);
	uint8_t* dValid = devAlloc<uint8_t>(keep, nbases);
	uint8_t* dOwn = devAlloc<uint8_t>(keep, nbases);
	uint8_t* dLater = devAlloc<uint8_t>(keep, nbases);
	uint32_t* dKeep = devAlloc<uint32_t>(keep, nbases);
	uint32_t* dSlot = devAlloc<uint32_t>(keep, nbases);
	uint32_t found = 0;
	if (nbases > 0)
	{
		GA_HIP(hipMemsetAsync(dLater, 0, nbases, st));
		hipLaunchKernelGGL(gad::idxHashKernel, dim3(blocksFor(nbases)), dim3(256), 0, st, dCode, nbases, k, dKmer, dHash, dValid);
		hipLaunchKernelGGL(gad::idxWindowKernel, dim3(blocksFor(nbases)), dim3(256), 0, st, dCode, dValid, dHash, nbases, k, w, dOwn, dLater);
		hipLaunchKernelGGL(gad::idxKeepKernel, dim3(blocksFor(nbases)), dim3(256), 0, st, dOwn, dLater, dSplit, dBufOff, dg.nodeOff, dRank, dMinStart, nbases, dKeep);
		GA_HIP(hipGetLastError());
		found = scanU32(st, dKeep, dSlot, nbases, keep);
	}
	uint32_t n2 = 1;
	while (n2 < found) n2 <<= 1;
	uint64_t* dKey = devAlloc<uint64_t>(keep, n2);
	uint32_t* dVal = devAlloc<uint32_t>(keep, n2);
	uint64_t* dPosFound = devAlloc<uint64_t>(keep, found);
	uint32_t groups = 0;
	std::vector<uint32_t> hist;
	uint64_t cap = 16;
	if (found > 0)
	{
		hipLaunchKernelGGL(gad::idxCompactKernel, dim3(blocksFor(nbases)), dim3(256), 0, st, dKeep, dSlot, dKmer, dSplit, dBufOff, nbases, dKey, dVal, dPosFound);
		if (n2 > found) hipLaunchKernelGGL(gad::padKeysKernel, dim3(blocksFor(n2 - found)), dim3(256), 0, st, dKey, dVal, found, n2);
		for (uint32_t size = 2; size <= n2; size <<= 1)
			for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
				hipLaunchKernelGGL(gad::bitonicStepKernel, dim3(blocksFor(n2)), dim3(256), 0, st, dKey, dVal, n2, size, stride);
		GA_HIP(hipGetLastError());
		uint32_t* dHead = devAlloc<uint32_t>(keep, found);
		uint32_t* dHeadScan = devAlloc<uint32_t>(keep, found);
		hipLaunchKernelGGL(gad::idxHeadsKernel, dim3(blocksFor(found)), dim3(256), 0, st, dKey, found, dHead);
		GA_HIP(hipGetLastError());
		groups = scanU32(st, dHead, dHeadScan, found, keep);
		uint32_t* dGroupStart = devAlloc<uint32_t>(keep, groups);
		uint32_t* dFirst = devAlloc<uint32_t>(keep, found);
		GA_HIP(hipMemsetAsync(dFirst, 0, found * sizeof(uint32_t), st));
		hipLaunchKernelGGL(gad::idxGroupsKernel, dim3(blocksFor(found)), dim3(256), 0, st, dHead, dHeadScan, dVal, found, dGroupStart, dFirst);
		GA_HIP(hipGetLastError());
		scanU32(st, dFirst, dFirst, found, keep);
		uint32_t* dGroupRank = devAlloc<uint32_t>(keep, groups);
		uint32_t* dCountByRank = devAlloc<uint32_t>(keep, groups);
		uint32_t* dMaxCount = devAlloc<uint32_t>(keep, 1);
		GA_HIP(hipMemsetAsync(dMaxCount, 0, sizeof(uint32_t), st));
		hipLaunchKernelGGL(gad::idxGroupRankKernel, dim3(blocksFor(groups)), dim3(256), 0, st, dGroupStart, groups, found, dVal, dFirst, dGroupRank, dCountByRank, dMaxCount);
		GA_HIP(hipGetLastError());
		scanU32(st, dCountByRank, dCountByRank, groups, keep);
		uint64_t* dPositions = devAlloc<uint64_t>(keep, found);
		hipLaunchKernelGGL(gad::idxPlaceKernel, dim3(blocksFor(found)), dim3(256), 0, st, dHead, dVal, dHeadScan, found, dGroupStart, groups, dGroupRank, dCountByRank, dPosFound, dPositions);
		GA_HIP(hipGetLastError());
		while (cap < 2 * (uint64_t)groups + 16) cap <<= 1;
		gad::SeedTableSlot* dTable = devAlloc<gad::SeedTableSlot>(keep, cap);
		GA_HIP(hipMemsetAsync(dTable, 0xff, cap * sizeof(gad::SeedTableSlot), st));
		uint32_t maxCount = 0;
		GA_HIP(hipMemcpyAsync(&maxCount, dMaxCount, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		GA_HIP(hipStreamSynchronize(st));
		const uint32_t histSize = maxCount + 1;
		uint32_t* dHist = devAlloc<uint32_t>(keep, histSize);
		GA_HIP(hipMemsetAsync(dHist, 0, histSize * sizeof(uint32_t), st));
		hipLaunchKernelGGL(gad::idxTableKernel, dim3(blocksFor(groups)), dim3(256), 0, st, dKey, dGroupStart, groups, found, dGroupRank, dCountByRank, dTable, cap - 1, dHist, histSize);
		GA_HIP(hipGetLastError());
		std::vector<gad::SeedTableSlot> table(cap);
		idx.positions.resize(found);
		hist.resize(histSize);
		GA_HIP(hipMemcpyAsync(table.data(), dTable, cap * sizeof(gad::SeedTableSlot), hipMemcpyDeviceToHost, st));
		GA_HIP(hipMemcpyAsync(idx.positions.data(), dPositions, found * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
		GA_HIP(hipMemcpyAsync(hist.data(), dHist, histSize * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		GA_HIP(hipStreamSynchronize(st));
		idx.table.resize(cap);
		for (uint64_t i = 0; i < cap; i++)
			idx.table[i] = table[i].kmer == ~0ull ? KmerSlot{~0ull, 0, 0} : KmerSlot{table[i].kmer, table[i].start, table[i].count};
	}
	else
	{
		idx.table.assign(cap, KmerSlot{~0ull, 0, 0});
	}
	idx.mask = cap - 1;
	idx.numKmers = groups;
	// maxCount: the sorted counts without their smallest entry (the key the
	// reference skips is substituted as in buildMinimizerIndex), the one at the
	// kept fraction, plus one
	idx.maxCount = 0;
	if (groups > 1)
	{
		const uint64_t size = groups - 1;
		uint64_t index = (uint64_t)(size * (1.0 - discardFraction));
		if (index == size) index = size - 1;
		// rank index + 1 among all counts (the smallest one removed)
		uint64_t want = index + 1, seen = 0;
		for (uint32_t c = 0; c < hist.size(); c++)
		{
			seen += hist[c];
			if (seen > want)
			{
				idx.maxCount = (uint64_t)c + 1;
				break;
			}
		}
	}
	return idx;
}

} // namespace ga
