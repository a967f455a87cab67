// Host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "ga_device_types.h"
#include "ga_order_core.h"

namespace gad {

// One thread per request, lanesPerWave (1..64) requests per wave, compiled for
// at least wavesPerSimd (1, 4 or 8) resident waves per SIMD; arenaOffsets[i]
// locates request i's workspace. The single-lane kernel (lanesPerWave 1) holds
// at most capWaves of its workgroups per SIMD (0: no cap).
hipError_t launchExtend(hipStream_t stream, uint32_t nreq, uint32_t lanesPerWave, uint32_t wavesPerSimd, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, const uint64_t* arenaOffsets, const ArenaCaps& caps, SliceSummary* sums, const uint64_t* sumOffsets, const uint64_t* planes, uint32_t capWaves = kSingleLaneWavesPerSimd);

// One read's seed loop per wave (lane 0 unless lanesPerWave > 1): jobs[i] locates
// read i's bytes, seeds and workspaces; results[i] receives its outcome.
hipError_t launchReadLoops(hipStream_t stream, uint32_t njobs, uint32_t lanesPerWave, uint32_t wavesPerSimd, const DevGraph& g, const ReadLoopParams& p, const ReadJob* jobs, const char* reads, const DevSeed* seeds, uint64_t* rawTraces, uint64_t* pools, DevAlignment* alns, ReadResult* results, unsigned char* arenas, const ArenaCaps& caps, const uint64_t* planes);

// Batched persistent form of launchReadLoops: zeroes *counter, then `workers`
// single-lane workers pull jobs [0, *njobs) in order from it. Worker w uses
// arenas + w * arenaStride and rawTraces + w * rawStride (two trace regions of
// each job's traceCap items), which must fit the batch's longest read.
hipError_t launchReadLoopsPersistent(hipStream_t stream, uint32_t workers, const DevGraph& g, const ReadLoopParams& p, const ReadJob* jobs, const uint32_t* njobs, uint32_t* counter, const char* reads, const DevSeed* seeds, uint64_t* rawTraces, uint64_t rawStride, uint64_t* pools, DevAlignment* alns, ReadResult* results, unsigned char* arenas, uint64_t arenaStride, const ArenaCaps& caps, const uint64_t* planes);

// Match-mask planes (planeBlock) of bytes[0, n): planeBlockCount(n) blocks of
// kPlaneWords words. The extension kernels take planes of their sequence buffer.
hipError_t launchEqPlanes(hipStream_t stream, const char* bytes, uint64_t n, uint64_t* planes);
uint64_t eqPlaneWords(uint64_t n);

// Workspace bytes of one request of length len (capacities for slices and band
// items scale with len; the others come from caps).
uint64_t requestArenaBytes(uint32_t len, const ArenaCaps& caps);

hipError_t launchSeedLookup(hipStream_t stream, const char* reads, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, const SeedTableSlot* table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* out, uint32_t* outCount);

// Copies matches[readOff[i], + count[i]) of every read i to out[outOff[i], ...).
// seed expansion and OrderSeeds of nreads reads (ga_order_core.h), one thread each
hipError_t launchSeedOrder(hipStream_t stream, const SeedGraphView& g, const uint64_t* positions, uint64_t npositions, uint64_t maxCount, uint32_t k, SeedMatch* matches,
                           const uint64_t* matchOff, const uint32_t* matchCount, const uint32_t* readLen, uint32_t nreads, double density, DevSeedHit* seeds,
                           ChainKey* scratch, const uint64_t* seedOff, const uint64_t* seedCap, int64_t* nseeds);
hipError_t launchGatherMatches(hipStream_t stream, const SeedMatch* matches, const uint64_t* readOff, const uint32_t* count, const uint64_t* outOff, uint32_t nreads, SeedMatch* out);

// Copies src[srcOff[i], + count[i]) to dst[dstOff[i], ...) for every range i.
hipError_t launchGatherU64(hipStream_t stream, const uint64_t* src, const uint64_t* srcOff, const uint32_t* count, const uint64_t* dstOff, uint32_t nranges, uint64_t* dst);

// Compact copy of the trace ranges src[srcOff[i], + count[i]): one code byte per
// item (ga_trace_codes.h) to codes[dstOff[i], ...); every escaped item is
// appended to escapes as the pair (index into codes, packed item), counted in
// *escapeCount (which the caller zeroes; escapeCap pairs fit).
hipError_t launchTraceCodes(hipStream_t stream, const DevGraph& g, const uint64_t* src, const uint64_t* srcOff, const uint32_t* count, const uint64_t* dstOff, uint32_t nranges, uint8_t* codes, uint64_t* escapes, uint32_t* escapeCount, uint64_t escapeCap);

// Device scratch of the allocation-free extension entry point: one file-scope
// array per GPU (the library holds it on every GPU it runs on until the process
// ends). noAllocScratchAddress returns this device's copy of it; call it with
// `device` current. It looks the address up once per device and keeps it.
constexpr uint64_t kNoAllocScratchBytes = 128ull << 20;
hipError_t noAllocScratchAddress(int device, void** ptr);

hipError_t launchRevcomp(hipStream_t stream, const char* reads, char* rev, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, uint32_t maxLen);

} // namespace gad
