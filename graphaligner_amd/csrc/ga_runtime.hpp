// Device runtime: one context per GPU (stream, resident graph copies, growable
// device buffers) and the batched extension driver with overflow reruns.
#pragma once
#include <hip/hip_runtime.h>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#include "ga_device_types.h"
#include "ga_backend.hpp"
#include "ga_graph.hpp"
#include "ga_index.hpp"

namespace ga {

struct HipError : std::runtime_error
{
	using std::runtime_error::runtime_error;
};

#define GA_HIP(x) \
	do { \
		hipError_t e_ = (x); \
		if (e_ != hipSuccess) throw ::ga::HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
	} while (0)

struct DeviceBuffer
{
	void* ptr = nullptr;
	size_t bytes = 0;
	DeviceBuffer() = default;
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
	~DeviceBuffer();
	void reserve(size_t b); // grows, never shrinks; contents are not kept
	template <typename T>
	T* as() const { return (T*)ptr; }
};

// Page-locked host memory (hipHostMalloc) for device-to-host copies at the full
// link rate.
struct PinnedBuffer
{
	void* ptr = nullptr;
	size_t bytes = 0;
	PinnedBuffer() = default;
	PinnedBuffer(const PinnedBuffer&) = delete;
	PinnedBuffer& operator=(const PinnedBuffer&) = delete;
	~PinnedBuffer();
	void reserve(size_t b); // grows, never shrinks; contents are not kept
	template <typename T>
	T* as() const { return (T*)ptr; }
};

// Stream, events and workspaces of one extension pipeline: extend calls on
// different lanes may run at the same time from different host threads, and
// their launches overlap on the GPU.
struct ExtendLane
{
	hipStream_t stream = nullptr;
	// single-lane extension workgroups per SIMD (launchExtend's capWaves)
	uint32_t capWaves = gad::kSingleLaneWavesPerSimd;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	DeviceBuffer arena, arenaOffsets, reqs, seqs, traces, results, sums, sumOffsets;
	DeviceBuffer gatherSrc, gatherDst, gatherCnt, traceCodes, traceEscapes, escapeCount, planes, bands;
	PinnedBuffer hostCodes; // compact traces of the lane's current extend call
	uint64_t hostCodesUsed = 0;
	ExtendLane() = default;
	ExtendLane(const ExtendLane&) = delete;
	ExtendLane& operator=(const ExtendLane&) = delete;
	~ExtendLane();
};

struct DeviceGraph
{
	uint64_t uid = 0;
	DeviceBuffer nodeLen, seq2, amb4, comp, lin, inOff, inAdj, outOff, outAdj, nodeId, nodeOff, bigSize;
	// what seed expansion and OrderSeeds read besides nodeId / nodeOff
	DeviceBuffer reverse, chain, chainPos;
	DeviceBuffer rec, outEdges;
	gad::DevGraph view;
	gad::SeedGraphView seedView;
};

struct DeviceIndex
{
	uint64_t uid = 0;
	DeviceBuffer table;
	uint64_t mask = 0;
	DeviceBuffer positions; // for seed expansion on the device
	uint64_t npositions = 0;
};

class Context : public ExtendBackend
{
public:
	explicit Context(int device);
	~Context();
	int device() const { return device_; }
	int computeUnits() const { return cus_; }
	hipStream_t stream() const { return stream_; }
	const gad::DevGraph& deviceGraph(const GraphHost& g);
	void extend(const GraphHost& g, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, bool wantSummaries, ExtBatch& out, const char* devSeqBase = nullptr, uint32_t lane = 0) override;
	uint32_t extendLanes() const override { return (uint32_t)lanes_.size(); }
	const DeviceIndex& deviceIndex(const MinimizerIndex& idx);
	void uploadReads(const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens) override;
	const ReadSet& reads() const override { return reads_; }
	void lookupSeeds(const MinimizerIndex& idx, uint32_t first, uint32_t count, SeedMatches& out) override;
	void orderedSeeds(const GraphHost& g, const MinimizerIndex& idx, uint32_t first, uint32_t count, const SeedMatches& sm, double density,
	                  std::vector<std::vector<SeedHit>>& seeds, std::vector<uint8_t>& asserted) override;
	const DeviceGraph& deviceGraphBuffers(const GraphHost& g);
	void prepareReads(uint32_t first, uint32_t count) override;
	void beginAlignmentCall() override;
	const char* readsBase() const override { return readBytes_.as<char>(); }
	void runReadLoops(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out) override;
	// The batched persistent form: the same read loops (same outputs) run by a
	// fixed set of workers pulling reads from a device counter, one captured
	// graph replayed per batch; the worker count is chosen by timing candidates
	// on a sample of the first batch. Reads whose workspaces overflow are rerun
	// by runReadLoops.
	void runReadLoopsBatched(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out) override;
	// The graph's device copy if it is resident on this context, else null
	// (never uploads).
	const gad::DevGraph* residentGraph(const GraphHost& g);
	// Extensions of ga_extend_batch_noalloc: n requests (split node and offset)
	// in the file-scope device scratch, with no device or host allocation. Traces
	// go to traces + 4 * i * traceCap as rows (split node, offset, seqPos, switch).
	// Returns false, touching nothing, when the batch does not fit the scratch
	// (kNoAllocMaxRequests requests, kNoAllocScratchBytes bytes); results of
	// extensions that still overflow after the reruns the scratch allows keep
	// status EXT_OVERFLOW.
	bool extendNoAlloc(const gad::DevGraph& dg, const gad::ExtParams& p, uint32_t n, const char* const* seqs, const uint32_t* lens, const uint32_t* startNodes, const uint32_t* startOffsets, int64_t traceCap, int64_t* traces, gad::ExtResult* results, double* kernelMs);
	static constexpr uint32_t kNoAllocMaxRequests = 4096;
	uint64_t reverseBase() const override { return revBase_; }
	// device workspace of the launches in flight (MI355X has 288 GB of HBM): a
	// device seed-loop launch may use all of it, an extension launch its lane's
	// share
	uint64_t arenaBudgetBytes = 160ull << 30;
	// Extension requests per wave: the GA_LANES_PER_WAVE environment variable
	// (1..64) if set, else defaultLanesPerWave. One request per wave measured
	// fastest on gfx950 even for 40,000-request launches: the requests of one
	// wave diverge almost completely, so its lanes run nearly one after another.
	uint32_t lanesPerWave(uint32_t nreq) const;
	uint32_t defaultLanesPerWave = 1;
	// register-budget variant of the extension kernel (resident waves per SIMD it
	// is compiled for: 1, 4 or 8); GA_KERNEL_WAVES_PER_SIMD overrides it
	uint32_t kernelWavesPerSimd = 1;

private:
	void runOnce(ExtendLane& L, const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out, const char* devSeqBase, const uint64_t* devPlanes);
	void runReadLoopLaunch(const gad::DevGraph& dg, const gad::ReadLoopParams& p, const std::vector<ReadLoopJob>& jobs, const std::vector<uint32_t>& which, const ReadLoopCaps& caps, ReadLoopOutput& out);
	// copies the finished reads' alignment records and used pool prefixes of one
	// launch (jobs rj, results res of jobs `which`) into out
	void collectReadLoops(const std::vector<gad::ReadJob>& rj, const std::vector<gad::ReadResult>& res, const std::vector<uint32_t>& which, ReadLoopOutput& out, const uint64_t* alnSrc, const uint64_t* poolSrc);
	// one persistent batch with `workers` workers (graph captured on first use
	// of a launch shape); returns the kernel ms
	float runPersistentBatch(const gad::DevGraph& dg, const gad::ReadLoopParams& p, const std::vector<gad::ReadJob>& rj, uint32_t workers, const ReadLoopCaps& caps, uint32_t maxLen, std::vector<gad::ReadResult>& res);
	void runLaunch(ExtendLane& L, const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out, const char* devSeqBase, const uint64_t* devPlanes);
	// match-mask planes of the resident reads (forward and reverse bytes);
	// computePlanes runs with cacheMutex_ held
	const uint64_t* residentPlanes(const char* devSeqBase);
	void computePlanes();
	// reverse complements of resident reads [first, first + count) unless already
	// written, then the planes; call with the lookup's events recorded around it
	void writeReverse(uint32_t first, uint32_t count);
	int device_;
	int cus_ = 0;
	hipStream_t stream_ = nullptr;
	hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
	// device copies of the most recently used graphs and indexes, by uid
	static constexpr size_t kCachedObjects = 2;
	std::vector<std::unique_ptr<DeviceGraph>> graphs_;
	std::vector<std::unique_ptr<DeviceIndex>> indexes_;
	static constexpr size_t kExtendLanes = 4;
	std::vector<std::unique_ptr<ExtendLane>> lanes_;
	std::mutex cacheMutex_; // graph and index caches, shared by the lanes
	DeviceBuffer arena_, gatherSrc_, gatherDst_, gatherCnt_, traceOut_; // device seed loop
	DeviceBuffer rlSeeds_, rlJobs_, rlTraces_, rlPool_, rlAlns_, rlResults_;
	DeviceBuffer readBytes_, readOff_, readLen_, matches_, matchCount_, matchOff_, gathered_, readPlanes_;
	DeviceBuffer seedOut_, seedScratch_, seedOff_, seedCap_, seedCount_;
	uint32_t lookupFirst_ = 0, lookupCount_ = 0; // read range of the last lookupSeeds
	// host staging of a lookup's matches (SeedMatches::staged)
	PinnedBuffer matchesHost_;
	// resident reads whose reverse complements (and so the planes) are written:
	// [revLo_, revHi_)
	uint64_t revLo_ = 0, revHi_ = 0;
	bool planesReady_ = false;
	ReadSet reads_;
	uint64_t revBase_ = 0;
	uint32_t lanesOverride_ = 0;
	// this device's copy of the no-allocation scratch (looked up at construction)
	unsigned char* scratch_ = nullptr;
	// batched persistent form: buffers, the captured graph of the current launch
	// shape, and the autotuned worker count (0: not tuned yet)
	DeviceBuffer bJobs_, bCounter_, bTraces_, bArena_, bPool_, bAlns_, bResults_;
	struct PersistentShape
	{
		const void* ptrs[9];
		uint32_t workers;
		uint64_t rawStride, arenaStride;
		gad::ArenaCaps caps;
		gad::ReadLoopParams p;
		gad::DevGraph dg;
	};
	PersistentShape graphShape_{};
	hipGraph_t graph_ = nullptr;
	hipGraphExec_t graphExec_ = nullptr;
public:
	uint32_t batchedWorkers = 0;
	double batchedTuneMs = 0;
	uint64_t graphLaunches = 0, graphCaptures = 0;
};

} // namespace ga
