// Device runtime: one context per GPU (stream, resident graph copies, growable
// device buffers) and the batched extension driver with overflow reruns.
#pragma once
#include <hip/hip_runtime.h>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#include "ga_device_types.h"
#include "ga_graph.hpp"

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

struct DeviceGraph
{
	DeviceBuffer nodeLen, seq2, amb4, comp, lin, inOff, inAdj, outOff, outAdj;
	gad::DevGraph view;
};

// One extension: align seq[0, len) from column startOffset of split node startNode.
struct ExtJob
{
	const char* seq;
	uint32_t len;
	uint32_t startNode;
	uint32_t startOffset;
};

struct ExtBatch
{
	std::vector<gad::ExtResult> results;
	std::vector<uint64_t> traceOffset; // per job, into trace
	std::vector<uint64_t> trace;       // packed items, each job's trace end first
	std::vector<gad::SliceSummary> summaries;
	std::vector<uint64_t> summaryOffset;
	double kernelMs = 0;               // summed kernel time of all launches
	uint32_t launches = 0;
	uint64_t wordSteps = 0;
};

class Context
{
public:
	explicit Context(int device);
	~Context();
	int device() const { return device_; }
	int computeUnits() const { return cus_; }
	hipStream_t stream() const { return stream_; }
	const gad::DevGraph& deviceGraph(const GraphHost& g);
	// Runs all jobs; statuses EXT_OVERFLOW are rerun with larger workspaces and
	// only remain if the largest workspace still does not fit.
	void extend(const GraphHost& g, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, bool wantSummaries, ExtBatch& out);
	uint64_t arenaBudgetBytes = 32ull << 30;

private:
	void runOnce(const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out);
	void runLaunch(const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out);
	int device_;
	int cus_ = 0;
	hipStream_t stream_ = nullptr;
	hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
	std::unordered_map<const GraphHost*, std::unique_ptr<DeviceGraph>> graphs_;
	DeviceBuffer arena_, arenaOffsets_, reqs_, seqs_, traces_, results_, sums_, sumOffsets_;
};

} // namespace ga
