// gfx950 kernels: seed extension with one extension per thread. Requests are
// sorted longest first so the threads of a wave carry similar lengths; a wave
// runs lanesPerWave requests (its other lanes exit at once), so a launch with
// few requests spreads them over many SIMDs instead of serialising them in the
// divergent lanes of a few waves.
#define GA_THREAD_PER_EXTENSION 1
#include <hip/hip_runtime.h>
#include <cstdlib>
#include <mutex>
#include "ga_extend_core.h"
#include "ga_read_core.h"
#include "ga_kernels.hpp"
#include "ga_seed_core.h"
#include "ga_order_core.h"
#include "ga_trace_codes.h"

namespace gad {

// Scratch of the allocation-free extension entry point (requests, results,
// sequences, match-mask planes, traces and workspaces of one call). In host code
// the name stands for no device address: host code uses only the address that
// noAllocScratchAddress looked up.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noAllocScratchAddress(int device, void** ptr)
{
	constexpr int kMaxDevices = 64;
	static std::mutex m;
	static void* addr[kMaxDevices] = {};
	if (device < 0 || device >= kMaxDevices || !ptr) return hipErrorInvalidValue;
	std::lock_guard<std::mutex> lock(m);
	if (!addr[device])
	{
		int current = -1;
		hipError_t e = hipGetDevice(&current);
		if (e != hipSuccess) return e;
		if (current != device) return hipErrorInvalidDevice;
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		if (!p) return hipErrorInvalidSymbol;
		addr[device] = p;
	}
	*ptr = addr[device];
	return hipSuccess;
}

template <int MinWavesPerSimd>
__global__ void __launch_bounds__(64, MinWavesPerSimd) extendThreadKernel(DevGraph g, ExtParams p, const ExtRequest* __restrict__ reqs, uint32_t nreq, uint32_t lanesPerWave, const char* __restrict__ seqbuf, uint64_t* __restrict__ tracebuf, ExtResult* __restrict__ results, unsigned char* __restrict__ arenas, const uint64_t* __restrict__ arenaOffsets, ArenaCaps caps, SliceSummary* __restrict__ sums, const uint64_t* __restrict__ sumOffsets, const uint64_t* __restrict__ planes)
{
	if (threadIdx.x >= lanesPerWave) return;
	uint32_t idx = blockIdx.x * lanesPerWave + threadIdx.x;
	if (idx >= nreq) return;
	ExtRequest r = reqs[idx];
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(r.len);
	c.maxItems = requestItems(r.len, caps);
	ExtResult out;
	runRequest(g, p, r, seqbuf, tracebuf, arenas + arenaOffsets[idx], c, sums ? sums + sumOffsets[idx] : nullptr, out, nullptr, planes);
	results[idx] = out;
}

// One request per single-thread workgroup: the compiler knows a single lane
// runs, so every value is wavefront-uniform and the extension's control flow and
// most of its 64-bit word arithmetic run on the scalar unit, with no exec-mask
// handling for its data-dependent branches. LdsQueue: the band queue lives in
// LDS (first attempts, whose capacities fit it).
template <bool LdsQueue>
__global__ void __launch_bounds__(1) extendSingleLaneKernel(DevGraph g, ExtParams p, const ExtRequest* __restrict__ reqs, uint32_t nreq, const char* __restrict__ seqbuf, uint64_t* __restrict__ tracebuf, ExtResult* __restrict__ results, unsigned char* __restrict__ arenas, const uint64_t* __restrict__ arenaOffsets, ArenaCaps caps, SliceSummary* __restrict__ sums, const uint64_t* __restrict__ sumOffsets, const uint64_t* __restrict__ planes)
{
	// the forward pass's queue and items and the traceback's columns share it
	__shared__ uint64_t ldsWords[LdsQueue ? kQueueWords : kColumnWords];
	uint32_t idx = blockIdx.x;
	if (idx >= nreq) return;
	ExtRequest r = reqs[idx];
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(r.len);
	c.maxItems = requestItems(r.len, caps);
	ExtResult out;
	const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
	runRequest(g, p, r, seqbuf, tracebuf, arenas + arenaOffsets[idx], c, sums ? sums + sumOffsets[idx] : nullptr, out, ldsWords, planes, true, ldsWords, LdsQueue);
	out.clockStart = t0;
	out.clockEnd = __builtin_amdgcn_s_memrealtime();
	results[idx] = out;
}

// LDS that a single-lane workgroup reserves beyond its own so that at most
// `waves` of them share a SIMD (0: no cap; GA_EXT_WAVES_PER_SIMD overrides). A
// launch bound by its longest extension lasts longer the more waves share that
// extension's SIMD issue slots; kSingleLaneWavesPerSimd (five) was the best
// measured setting on MI355X for the 5 Mbp variation-graph workload (the kernel
// needs few enough registers for seven).
static uint32_t singleLaneExtraLds(uint32_t own, uint32_t waves)
{
	static const uint32_t ldsPerCu = []() -> uint32_t {
		int dev = 0, v = 0;
		if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess) return 0;
		return (uint32_t)v;
	}();
	static const int envWaves = std::getenv("GA_EXT_WAVES_PER_SIMD") ? (int)std::strtoul(std::getenv("GA_EXT_WAVES_PER_SIMD"), nullptr, 10) : -1;
	if (envWaves >= 0) waves = (uint32_t)envWaves;
	if (waves == 0 || ldsPerCu == 0) return 0;
	// four SIMDs per CU; whole 512-byte LDS allocation blocks
	const uint32_t perGroup = (ldsPerCu / (4 * waves)) & ~511u;
	return perGroup > own ? perGroup - own : 0;
}

hipError_t launchExtend(hipStream_t stream, uint32_t nreq, uint32_t lanesPerWave, uint32_t wavesPerSimd, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, const uint64_t* arenaOffsets, const ArenaCaps& caps, SliceSummary* sums, const uint64_t* sumOffsets, const uint64_t* planes, uint32_t capWaves)
{
	if (lanesPerWave < 1 || lanesPerWave > 64) return hipErrorInvalidValue;
	if (lanesPerWave == 1 && wavesPerSimd <= 1)
	{
		if (queueFitsLds(caps))
			hipLaunchKernelGGL(extendSingleLaneKernel<true>, dim3(nreq), dim3(1), singleLaneExtraLds(kQueueWords * 8, capWaves), stream, g, p, reqs, nreq, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets, planes);
		else
			hipLaunchKernelGGL(extendSingleLaneKernel<false>, dim3(nreq), dim3(1), singleLaneExtraLds(kColumnWords * 8, capWaves), stream, g, p, reqs, nreq, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets, planes);
		return hipGetLastError();
	}
	uint32_t blocks = (nreq + lanesPerWave - 1) / lanesPerWave;
	// register budget variants: more resident waves per SIMD for fewer VGPRs
	if (wavesPerSimd >= 8)
		hipLaunchKernelGGL(extendThreadKernel<8>, dim3(blocks), dim3(64), 0, stream, g, p, reqs, nreq, lanesPerWave, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets, planes);
	else if (wavesPerSimd >= 4)
		hipLaunchKernelGGL(extendThreadKernel<4>, dim3(blocks), dim3(64), 0, stream, g, p, reqs, nreq, lanesPerWave, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets, planes);
	else
		hipLaunchKernelGGL(extendThreadKernel<1>, dim3(blocks), dim3(64), 0, stream, g, p, reqs, nreq, lanesPerWave, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets, planes);
	return hipGetLastError();
}

// One read's whole seed loop per wave (lane 0): extensions of one read follow
// each other with no host round trip, and reads are independent, so the
// hardware dispatcher balances the uneven per-read work across the chip.
template <int MinWavesPerSimd>
__global__ void __launch_bounds__(64, MinWavesPerSimd) readLoopKernel(DevGraph g, ReadLoopParams p, const ReadJob* __restrict__ jobs, uint32_t njobs, uint32_t lanesPerWave, const char* __restrict__ reads, const DevSeed* __restrict__ seeds, uint64_t* __restrict__ rawTraces, uint64_t* __restrict__ pools, DevAlignment* __restrict__ alns, ReadResult* __restrict__ results, unsigned char* __restrict__ arenas, ArenaCaps caps, const uint64_t* __restrict__ planes)
{
	if (threadIdx.x >= lanesPerWave) return;
	uint32_t idx = blockIdx.x * lanesPerWave + threadIdx.x;
	if (idx >= njobs) return;
	const ReadJob j = jobs[idx];
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(j.len);
	c.maxItems = requestItems(j.len, caps);
	ReadResult out;
	runReadLoop(g, p, j, reads, seeds + j.seedStart, rawTraces + j.traceOffset, pools + j.poolOffset, alns + j.alnOffset, arenas + j.arenaOffset, c, out, nullptr, planes);
	results[idx] = out;
}

__global__ void __launch_bounds__(1) readLoopSingleLaneKernel(DevGraph g, ReadLoopParams p, const ReadJob* __restrict__ jobs, uint32_t njobs, const char* __restrict__ reads, const DevSeed* __restrict__ seeds, uint64_t* __restrict__ rawTraces, uint64_t* __restrict__ pools, DevAlignment* __restrict__ alns, ReadResult* __restrict__ results, unsigned char* __restrict__ arenas, ArenaCaps caps, const uint64_t* __restrict__ planes)
{
	__shared__ uint64_t colStore[kColumnWords];
	uint32_t idx = blockIdx.x;
	if (idx >= njobs) return;
	const ReadJob j = jobs[idx];
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(j.len);
	c.maxItems = requestItems(j.len, caps);
	ReadResult out;
	runReadLoop(g, p, j, reads, seeds + j.seedStart, rawTraces + j.traceOffset, pools + j.poolOffset, alns + j.alnOffset, arenas + j.arenaOffset, c, out, colStore, planes, true);
	results[idx] = out;
}

// Batched persistent form: a fixed set of single-lane workers, each pulling whole
// reads (in the order of `jobs`, longest first) from a device counter until the
// batch is drained; a worker reuses its own workspace and raw-trace region for
// every read it takes, so memory scales with the workers, not the reads. The
// number of jobs is read from device memory so that one captured graph serves
// every batch. Every worker exits once the counter passes the job count.
template <bool LdsQueue>
__global__ void __launch_bounds__(1) readLoopPersistentKernel(DevGraph g, ReadLoopParams p, const ReadJob* __restrict__ jobs, const uint32_t* __restrict__ njobs, uint32_t* __restrict__ next, const char* __restrict__ reads, const DevSeed* __restrict__ seeds, uint64_t* __restrict__ rawTraces, uint64_t rawStride, uint64_t* __restrict__ pools, DevAlignment* __restrict__ alns, ReadResult* __restrict__ results, unsigned char* __restrict__ arenas, uint64_t arenaStride, ArenaCaps caps, const uint64_t* __restrict__ planes)
{
	__shared__ uint64_t ldsWords[LdsQueue ? kQueueWords : kColumnWords];
	const uint32_t n = uni32(*njobs);
	unsigned char* arena = arenas + (uint64_t)blockIdx.x * arenaStride;
	uint64_t* raw = rawTraces + (uint64_t)blockIdx.x * rawStride;
	while (true)
	{
		const uint32_t idx = uni32(atomicAdd(next, 1u));
		if (idx >= n) break;
		const ReadJob j = jobs[idx];
		ArenaCaps c = caps;
		c.maxSlices = requestSlices(j.len);
		c.maxItems = requestItems(j.len, caps);
		ReadResult out;
		runReadLoop(g, p, j, reads, seeds + j.seedStart, raw, pools + j.poolOffset, alns + j.alnOffset, arena, c, out, ldsWords, planes, true, ldsWords, LdsQueue);
		results[idx] = out;
	}
}

// Zeroes the batch counter (a kernel, so that it can sit in a captured graph).
__global__ void __launch_bounds__(64) resetCounterKernel(uint32_t* c)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

hipError_t launchReadLoopsPersistent(hipStream_t stream, uint32_t workers, const DevGraph& g, const ReadLoopParams& p, const ReadJob* jobs, const uint32_t* njobs, uint32_t* counter, const char* reads, const DevSeed* seeds, uint64_t* rawTraces, uint64_t rawStride, uint64_t* pools, DevAlignment* alns, ReadResult* results, unsigned char* arenas, uint64_t arenaStride, const ArenaCaps& caps, const uint64_t* planes)
{
	if (workers == 0) return hipErrorInvalidValue;
	hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(64), 0, stream, counter);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	if (queueFitsLds(caps))
		hipLaunchKernelGGL(readLoopPersistentKernel<true>, dim3(workers), dim3(1), 0, stream, g, p, jobs, njobs, counter, reads, seeds, rawTraces, rawStride, pools, alns, results, arenas, arenaStride, caps, planes);
	else
		hipLaunchKernelGGL(readLoopPersistentKernel<false>, dim3(workers), dim3(1), 0, stream, g, p, jobs, njobs, counter, reads, seeds, rawTraces, rawStride, pools, alns, results, arenas, arenaStride, caps, planes);
	return hipGetLastError();
}

hipError_t launchReadLoops(hipStream_t stream, uint32_t njobs, uint32_t lanesPerWave, uint32_t wavesPerSimd, const DevGraph& g, const ReadLoopParams& p, const ReadJob* jobs, const char* reads, const DevSeed* seeds, uint64_t* rawTraces, uint64_t* pools, DevAlignment* alns, ReadResult* results, unsigned char* arenas, const ArenaCaps& caps, const uint64_t* planes)
{
	if (lanesPerWave < 1 || lanesPerWave > 64) return hipErrorInvalidValue;
	if (lanesPerWave == 1 && wavesPerSimd <= 1)
	{
		hipLaunchKernelGGL(readLoopSingleLaneKernel, dim3(njobs), dim3(1), 0, stream, g, p, jobs, njobs, reads, seeds, rawTraces, pools, alns, results, arenas, caps, planes);
		return hipGetLastError();
	}
	uint32_t blocks = (njobs + lanesPerWave - 1) / lanesPerWave;
	if (wavesPerSimd >= 4)
		hipLaunchKernelGGL(readLoopKernel<4>, dim3(blocks), dim3(64), 0, stream, g, p, jobs, njobs, lanesPerWave, reads, seeds, rawTraces, pools, alns, results, arenas, caps, planes);
	else
		hipLaunchKernelGGL(readLoopKernel<1>, dim3(blocks), dim3(64), 0, stream, g, p, jobs, njobs, lanesPerWave, reads, seeds, rawTraces, pools, alns, results, arenas, caps, planes);
	return hipGetLastError();
}

// planeBlock with one lane per byte: each wavefront walks 64-position blocks
// (block, block + waves, ...), its byte loads contiguous, and a block's five
// words are the ballots of the lanes' letter sets (readBaseSet for all 256 byte
// values in LDS; positions at or past n set no bit).
__global__ void __launch_bounds__(256) eqPlanesKernel(const char* __restrict__ bytes, uint64_t n, uint64_t blocks, uint64_t* __restrict__ planes)
{
	__shared__ uint8_t sets[256];
	sets[threadIdx.x] = (uint8_t)readBaseSet((unsigned char)threadIdx.x);
	__syncthreads();
	const uint32_t lane = threadIdx.x & 63;
	const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
	for (uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; b < blocks; b += waves)
	{
		const uint64_t x = b * 64 + lane;
		const uint32_t set = x < n ? sets[(unsigned char)bytes[x]] : 0;
		const uint64_t m0 = __ballot((set & 1) != 0);
		const uint64_t m1 = __ballot((set & 2) != 0);
		const uint64_t m2 = __ballot((set & 4) != 0);
		const uint64_t m3 = __ballot((set & 8) != 0);
		const uint64_t m4 = __ballot(set == 16);
		if (lane < kPlaneWords)
			planes[b * kPlaneWords + lane] = lane == 0 ? m0 : lane == 1 ? m1 : lane == 2 ? m2 : lane == 3 ? m3 : m4;
	}
}

uint64_t eqPlaneWords(uint64_t n) { return planeBlockCount(n) * kPlaneWords; }

hipError_t launchEqPlanes(hipStream_t stream, const char* bytes, uint64_t n, uint64_t* planes)
{
	const uint64_t blocks = planeBlockCount(n);
	// four waves per 256-thread workgroup, at most 16 workgroups per CU's worth
	// of the 256 CUs; each wave walks its share of the 64-position blocks
	const uint64_t grid = (blocks + 3) / 4 < 256 * 16 ? (blocks + 3) / 4 : 256 * 16;
	hipLaunchKernelGGL(eqPlanesKernel, dim3((uint32_t)grid), dim3(256), 0, stream, bytes, n, blocks, planes);
	return hipGetLastError();
}

uint64_t requestArenaBytes(uint32_t len, const ArenaCaps& caps)
{
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(len);
	c.maxItems = requestItems(len, caps);
	return arenaBytes(c);
}

// One workgroup per trace range, one item per thread and step: the code of item i
// needs only items i - 1 and i, so the range is encoded in parallel.
__global__ void traceCodesKernel(TraceGraph tg, const uint64_t* __restrict__ src, const uint64_t* __restrict__ srcOff, const uint32_t* __restrict__ count, const uint64_t* __restrict__ dstOff, uint32_t nranges, uint8_t* __restrict__ codes, uint64_t* __restrict__ escapes, uint32_t* __restrict__ escapeCount, uint64_t escapeCap)
{
	uint32_t r = blockIdx.x;
	if (r >= nranges) return;
	const uint64_t* s = src + srcOff[r];
	const uint64_t d = dstOff[r];
	const uint32_t c = count[r];
	for (uint32_t i = threadIdx.x; i < c; i += blockDim.x)
	{
		const uint64_t cur = s[i];
		const uint32_t code = i == 0 ? kTraceEscape : traceCode(tg, s[i - 1], cur);
		codes[d + i] = (uint8_t)code;
		if (code == kTraceEscape)
		{
			uint32_t slot = atomicAdd(escapeCount, 1u);
			if (slot < escapeCap)
			{
				escapes[2 * (uint64_t)slot] = d + i;
				escapes[2 * (uint64_t)slot + 1] = cur;
			}
		}
	}
}

hipError_t launchTraceCodes(hipStream_t stream, const DevGraph& g, const uint64_t* src, const uint64_t* srcOff, const uint32_t* count, const uint64_t* dstOff, uint32_t nranges, uint8_t* codes, uint64_t* escapes, uint32_t* escapeCount, uint64_t escapeCap)
{
	TraceGraph tg{g.nodeLen, g.inOff, g.inAdj};
	hipLaunchKernelGGL(traceCodesKernel, dim3(nranges), dim3(256), 0, stream, tg, src, srcOff, count, dstOff, nranges, codes, escapes, escapeCount, escapeCap);
	return hipGetLastError();
}

} // namespace gad

namespace gad {

// base codes kept before each chunk of the read in seedLookupKernel: enough for
// the k - 1 positions before a k-mer's end (k <= kLookupHalo + 1)
constexpr uint32_t kLookupHalo = 32;

// One read per 256-thread workgroup. Positions go in chunks of 256, one per
// thread: each thread decides whether iterateKmers reports its k-mer, looks it
// up, and the workgroup writes the chunk's matches in read order (ballot prefix
// counts within each wavefront, wavefront totals through LDS), as
// lookupReadKmers writes them one by one. The rule of kmerReported (the k-mer
// ending at each position, reported unless it lies inside a run of one base
// longer than k off the run's window grid) evaluated from the chunk's base codes
// staged in LDS: every position reads its k-mer window in a loop of k steps
// without an early exit, and counts the run of its base within the window on the
// way, so the wave's lanes stay together; only a run filling the whole window
// continues backwards lane by lane, as kmerReported does.
__global__ void __launch_bounds__(256) seedLookupKernel(const char* __restrict__ reads, const uint64_t* __restrict__ readOff, const uint32_t* __restrict__ readLen, uint32_t nreads, const SeedTableSlot* __restrict__ table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* __restrict__ out, uint32_t* __restrict__ outCount)
{
	__shared__ uint32_t waveHits[4];
	__shared__ uint32_t written;
	__shared__ int8_t codes[kLookupHalo + 256];
	const uint32_t r = blockIdx.x;
	if (r >= nreads) return;
	const char* s = reads + readOff[r];
	const uint64_t len = readLen[r];
	SeedMatch* o = out + readOff[r];
	const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint64_t realWindow = w - k + 1;
	if (threadIdx.x == 0) written = 0;
	for (uint64_t c = 0; c < len; c += 256)
	{
		// the chunk's codes and the kLookupHalo positions before it (-1: not a
		// base, or before the read)
		const uint64_t i = c + threadIdx.x;
		codes[kLookupHalo + threadIdx.x] = i < len ? (int8_t)readBaseCode((unsigned char)s[i]) : (int8_t)-1;
		if (threadIdx.x < kLookupHalo)
		{
			const int64_t j = (int64_t)c - (int64_t)kLookupHalo + (int64_t)threadIdx.x;
			codes[threadIdx.x] = j >= 0 ? (int8_t)readBaseCode((unsigned char)s[j]) : (int8_t)-1;
		}
		__syncthreads();
		bool hit = false;
		SeedMatch m;
		m.pos = m.start = m.count = 0;
		if (i < len && i + 1 >= k)
		{
			const int32_t t = (int32_t)(kLookupHalo + threadIdx.x);
			const int32_t first = t - (int32_t)k + 1;
			const int32_t ci = codes[t];
			uint64_t kmer = 0;
			bool bad = false;
			// run of ci ending at i inside the window: positions i, i - 1, ...
			uint64_t run = 0;
			bool same = true;
			for (uint32_t q = 0; q < k; q++)
			{
				const int32_t cq = codes[first + (int32_t)q];
				bad |= cq < 0;
				kmer = (kmer << 2) | (uint64_t)(cq & 3);
				const int32_t cb = codes[t - (int32_t)q];
				same = same && cb == ci;
				run += same ? 1 : 0;
			}
			if (!bad)
			{
				uint64_t h = run;
				// a run filling the window may continue before it
				if (h == k)
					while (h <= i && readBaseCode((unsigned char)s[i - h]) == ci) h++;
				const bool reported = h <= k || (h - k) % realWindow == 0;
				if (reported)
				{
					const SeedTableSlot* ts = seedTableFind(table, mask, kmer);
					if (ts && ts->count < maxCount)
					{
						hit = true;
						m.pos = (uint32_t)i;
						m.start = ts->start;
						m.count = ts->count;
					}
				}
			}
		}
		const uint64_t ballot = __ballot(hit);
		const uint32_t before = (uint32_t)__popcll(ballot & (((uint64_t)1 << lane) - 1));
		if (lane == 0) waveHits[wave] = (uint32_t)__popcll(ballot);
		__syncthreads();
		uint32_t at = written;
		for (uint32_t v = 0; v < wave; v++) at += waveHits[v];
		if (hit) o[at + before] = m;
		__syncthreads();
		if (threadIdx.x == 0) written += waveHits[0] + waveHits[1] + waveHits[2] + waveHits[3];
		__syncthreads();
	}
	if (threadIdx.x == 0) outCount[r] = written;
}

// Reverse complement of every read, one block per read, through a table of
// complementByte for all 256 byte values in LDS (no per-letter branches).
__global__ void __launch_bounds__(256) revcompKernel(const char* __restrict__ reads, char* __restrict__ rev, const uint64_t* __restrict__ readOff, const uint32_t* __restrict__ readLen, uint32_t nreads)
{
	__shared__ char comp[256];
	comp[threadIdx.x] = complementByte((char)threadIdx.x);
	__syncthreads();
	uint32_t r = blockIdx.x;
	if (r >= nreads) return;
	uint64_t len = readLen[r];
	const char* s = reads + readOff[r];
	char* d = rev + readOff[r];
	for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) d[i] = comp[(unsigned char)s[len - 1 - i]];
}

hipError_t launchSeedLookup(hipStream_t stream, const char* reads, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, const SeedTableSlot* table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* out, uint32_t* outCount)
{
	// the k-mer window must lie in the chunk and its staged halo
	if (k == 0 || k > kLookupHalo + 1 || w < k) return hipErrorInvalidValue;
	hipLaunchKernelGGL(seedLookupKernel, dim3(nreads), dim3(256), 0, stream, reads, readOff, readLen, nreads, table, mask, k, w, maxCount, out, outCount);
	return hipGetLastError();
}

hipError_t launchRevcomp(hipStream_t stream, const char* reads, char* rev, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, uint32_t maxLen)
{
	(void)maxLen;
	hipLaunchKernelGGL(revcompKernel, dim3(nreads), dim3(256), 0, stream, reads, rev, readOff, readLen, nreads);
	return hipGetLastError();
}

} // namespace gad

namespace gad {

__global__ void gatherMatchesKernel(const SeedMatch* __restrict__ matches, const uint64_t* __restrict__ readOff, const uint32_t* __restrict__ count, const uint64_t* __restrict__ outOff, uint32_t nreads, SeedMatch* __restrict__ out)
{
	uint32_t r = blockIdx.x;
	if (r >= nreads) return;
	const SeedMatch* src = matches + readOff[r];
	SeedMatch* dst = out + outOff[r];
	uint32_t c = count[r];
	for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) dst[i] = src[i];
}

hipError_t launchGatherMatches(hipStream_t stream, const SeedMatch* matches, const uint64_t* readOff, const uint32_t* count, const uint64_t* outOff, uint32_t nreads, SeedMatch* out)
{
	hipLaunchKernelGGL(gatherMatchesKernel, dim3(nreads), dim3(256), 0, stream, matches, readOff, count, outOff, nreads, out);
	return hipGetLastError();
}

} // namespace gad

namespace gad {

__global__ void gatherU64Kernel(const uint64_t* __restrict__ src, const uint64_t* __restrict__ srcOff, const uint32_t* __restrict__ count, const uint64_t* __restrict__ dstOff, uint32_t nranges, uint64_t* __restrict__ dst)
{
	uint32_t r = blockIdx.x;
	if (r >= nranges) return;
	const uint64_t* s = src + srcOff[r];
	uint64_t* d = dst + dstOff[r];
	uint32_t c = count[r];
	for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) d[i] = s[i];
}

hipError_t launchGatherU64(hipStream_t stream, const uint64_t* src, const uint64_t* srcOff, const uint32_t* count, const uint64_t* dstOff, uint32_t nranges, uint64_t* dst)
{
	hipLaunchKernelGGL(gatherU64Kernel, dim3(nranges), dim3(256), 0, stream, src, srcOff, count, dstOff, nranges, dst);
	return hipGetLastError();
}

} // namespace gad

namespace gad {

// Seed expansion and OrderSeeds of one read per single-thread workgroup
// (seedsOfRead): each read's sorts are a serial chain, so a read alone in its
// wave costs that chain only, where 64 reads in one wave's lanes would diverge
// and run one after another; values are wave-uniform and go to the scalar unit.
// The read's matches are sorted in place by count, its seeds written at
// seedOff[r] with room for seedCap[r], and a same-sized scratch of chain keys
// used; the sorts' pending ranges are in LDS. nseeds[r] receives the count or a
// negative code (-1: a reference assertion, -2/-3: inconsistent input).
__global__ void __launch_bounds__(1) seedOrderKernel(SeedGraphView g, const uint64_t* __restrict__ positions, uint64_t npositions, uint64_t maxCount, uint32_t k,
                                                     SeedMatch* __restrict__ matches, const uint64_t* __restrict__ matchOff, const uint32_t* __restrict__ matchCount,
                                                     const uint32_t* __restrict__ readLen, uint32_t nreads, double density, DevSeedHit* __restrict__ seeds,
                                                     ChainKey* __restrict__ scratch, const uint64_t* __restrict__ seedOff, const uint64_t* __restrict__ seedCap,
                                                     int64_t* __restrict__ nseeds)
{
	__shared__ SortRange stack[kSortStackEntries];
	const uint32_t r = blockIdx.x;
	if (r >= nreads) return;
	nseeds[r] = seedsOfRead(g, positions, npositions, maxCount, k, matches + matchOff[r], matchCount[r], readLen[r], density, seeds + seedOff[r], seedCap[r],
	                        scratch + seedOff[r], stack);
}

hipError_t launchSeedOrder(hipStream_t stream, const SeedGraphView& g, const uint64_t* positions, uint64_t npositions, uint64_t maxCount, uint32_t k, SeedMatch* matches,
                           const uint64_t* matchOff, const uint32_t* matchCount, const uint32_t* readLen, uint32_t nreads, double density, DevSeedHit* seeds,
                           ChainKey* scratch, const uint64_t* seedOff, const uint64_t* seedCap, int64_t* nseeds)
{
	if (nreads == 0) return hipSuccess;
	hipLaunchKernelGGL(seedOrderKernel, dim3(nreads), dim3(1), 0, stream, g, positions, npositions, maxCount, k, matches, matchOff, matchCount, readLen, nreads,
	                   density, seeds, scratch, seedOff, seedCap, nseeds);
	return hipGetLastError();
}

} // namespace gad
