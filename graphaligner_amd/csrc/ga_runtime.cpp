// Device runtime (see ga_runtime.hpp).
#include "ga_runtime.hpp"
#include "ga_align.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include "ga_kernels.hpp"

namespace ga {

DeviceBuffer::~DeviceBuffer()
{
	if (ptr) (void)hipFree(ptr);
}

void DeviceBuffer::reserve(size_t b)
{
	if (b <= bytes && ptr) return;
	if (ptr) GA_HIP(hipFree(ptr));
	ptr = nullptr;
	bytes = 0;
	size_t want = std::max<size_t>(b, 256);
	GA_HIP(hipMalloc(&ptr, want));
	bytes = want;
}

PinnedBuffer::~PinnedBuffer()
{
	if (ptr) (void)hipHostFree(ptr);
}

void PinnedBuffer::reserve(size_t b)
{
	if (b <= bytes && ptr) return;
	if (ptr) GA_HIP(hipHostFree(ptr));
	ptr = nullptr;
	bytes = 0;
	size_t want = std::max<size_t>(b, 4096);
	GA_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
	bytes = want;
}

uint32_t Context::lanesPerWave(uint32_t nreq) const
{
	(void)nreq;
	return lanesOverride_ > 0 ? lanesOverride_ : defaultLanesPerWave;
}

Context::Context(int device) : device_(device)
{
	if (const char* v = std::getenv("GA_LANES_PER_WAVE"))
	{
		int x = std::atoi(v);
		if (x >= 1 && x <= 64) lanesOverride_ = (uint32_t)x;
	}
	if (const char* v = std::getenv("GA_KERNEL_WAVES_PER_SIMD"))
	{
		int x = std::atoi(v);
		if (x >= 1 && x <= 8) kernelWavesPerSimd = (uint32_t)x;
	}
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count == 0) throw HipError("no HIP device available: the GPU path needs an MI355X (gfx950)");
	if (device < 0 || device >= count) throw HipError("HIP device ordinal out of range");
	GA_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	GA_HIP(hipGetDeviceProperties(&prop, device));
	cus_ = prop.multiProcessorCount;
	GA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	GA_HIP(hipEventCreate(&ev0_));
	GA_HIP(hipEventCreate(&ev1_));
	void* sp = nullptr;
	GA_HIP(gad::noAllocScratchAddress(device, &sp));
	scratch_ = (unsigned char*)sp;
	for (size_t k = 0; k < kExtendLanes; k++)
	{
		auto L = std::make_unique<ExtendLane>();
		GA_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
		GA_HIP(hipEventCreate(&L->ev0));
		GA_HIP(hipEventCreate(&L->ev1));
		lanes_.push_back(std::move(L));
	}
}

ExtendLane::~ExtendLane()
{
	if (ev0) (void)hipEventDestroy(ev0);
	if (ev1) (void)hipEventDestroy(ev1);
	if (stream) (void)hipStreamDestroy(stream);
}

Context::~Context()
{
	(void)hipSetDevice(device_);
	lanes_.clear();
	graphs_.clear();
	indexes_.clear();
	if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
	if (graph_) (void)hipGraphDestroy(graph_);
	if (ev0_) (void)hipEventDestroy(ev0_);
	if (ev1_) (void)hipEventDestroy(ev1_);
	if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename T>
static void upload(DeviceBuffer& buf, const std::vector<T>& v, hipStream_t s)
{
	buf.reserve(std::max<size_t>(1, v.size()) * sizeof(T));
	if (!v.empty()) GA_HIP(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

const gad::DevGraph& Context::deviceGraph(const GraphHost& g) { return deviceGraphBuffers(g).view; }

const DeviceGraph& Context::deviceGraphBuffers(const GraphHost& g)
{
	if (g.uid == 0) throw std::runtime_error("graph has no identity (not built by loadGfaGraph)");
	std::lock_guard<std::mutex> lock(cacheMutex_);
	for (auto& c : graphs_)
		if (c->uid == g.uid) return *c;
	GA_HIP(hipSetDevice(device_));
	GA_HIP(hipStreamSynchronize(stream_));
	if (graphs_.size() >= kCachedObjects) graphs_.erase(graphs_.begin());
	auto dg = std::make_unique<DeviceGraph>();
	dg->uid = g.uid;
	upload(dg->nodeLen, g.nodeLen, stream_);
	upload(dg->seq2, g.seq2, stream_);
	upload(dg->amb4, g.amb4, stream_);
	upload(dg->comp, g.comp, stream_);
	upload(dg->lin, g.linearizable, stream_);
	upload(dg->inOff, g.inOff, stream_);
	upload(dg->inAdj, g.inAdj, stream_);
	upload(dg->outOff, g.outOff, stream_);
	upload(dg->outAdj, g.outAdj, stream_);
	upload(dg->nodeId, g.nodeId, stream_);
	upload(dg->nodeOff, g.nodeOff, stream_);
	std::vector<uint32_t> bigSize(g.numNodes());
	for (size_t i = 0; i < g.numNodes(); i++) bigSize[i] = (uint32_t)g.splitBigSize[i];
	upload(dg->bigSize, bigSize, stream_);
	upload(dg->reverse, g.reverse, stream_);
	upload(dg->chain, g.chain, stream_);
	upload(dg->chainPos, g.chainPos, stream_);
	upload(dg->rec, g.nodeRec, stream_);
	upload(dg->outEdges, g.outEdges, stream_);
	GA_HIP(hipStreamSynchronize(stream_));
	auto& v = dg->view;
	v.nodeLen = dg->nodeLen.as<uint32_t>();
	v.seq2 = dg->seq2.as<uint64_t>();
	v.amb4 = dg->amb4.as<uint64_t>();
	v.comp = dg->comp.as<uint32_t>();
	v.linearizable = dg->lin.as<uint8_t>();
	v.inOff = dg->inOff.as<uint32_t>();
	v.inAdj = dg->inAdj.as<uint32_t>();
	v.outOff = dg->outOff.as<uint32_t>();
	v.outAdj = dg->outAdj.as<uint32_t>();
	v.nodeId = dg->nodeId.as<int32_t>();
	v.nodeOff = dg->nodeOff.as<uint32_t>();
	v.bigSize = dg->bigSize.as<uint32_t>();
	v.rec = dg->rec.as<gad::NodeRec>();
	v.outEdges = dg->outEdges.as<gad::OutEdge>();
	v.numNodes = g.numNodes();
	v.firstAmbiguous = g.firstAmbiguous;
	dg->seedView = gad::SeedGraphView{v.nodeId, v.nodeOff, dg->reverse.as<uint8_t>(), dg->chain.as<uint32_t>(), dg->chainPos.as<uint64_t>()};
	graphs_.push_back(std::move(dg));
	return *graphs_.back();
}

const DeviceIndex& Context::deviceIndex(const MinimizerIndex& idx)
{
	if (idx.uid == 0) throw std::runtime_error("index has no identity (not built by buildMinimizerIndex)");
	std::lock_guard<std::mutex> lock(cacheMutex_);
	for (auto& c : indexes_)
		if (c->uid == idx.uid) return *c;
	GA_HIP(hipSetDevice(device_));
	GA_HIP(hipStreamSynchronize(stream_));
	if (indexes_.size() >= kCachedObjects) indexes_.erase(indexes_.begin());
	static_assert(sizeof(KmerSlot) == sizeof(gad::SeedTableSlot), "table slot layouts must agree");
	auto di = std::make_unique<DeviceIndex>();
	di->uid = idx.uid;
	upload(di->table, idx.table, stream_);
	upload(di->positions, idx.positions, stream_);
	GA_HIP(hipStreamSynchronize(stream_));
	di->mask = idx.mask;
	di->npositions = idx.positions.size();
	indexes_.push_back(std::move(di));
	return *indexes_.back();
}

void Context::computePlanes()
{
	const uint64_t n = 2 * (reads_.bytes + 1);
	readPlanes_.reserve(gad::eqPlaneWords(n) * sizeof(uint64_t));
	GA_HIP(gad::launchEqPlanes(stream_, readBytes_.as<char>(), n, readPlanes_.as<uint64_t>()));
	planesReady_ = true;
}

const uint64_t* Context::residentPlanes(const char* devSeqBase)
{
	if (devSeqBase != readBytes_.as<char>()) throw std::runtime_error("device sequences are not the context's resident reads");
	std::lock_guard<std::mutex> lock(cacheMutex_);
	if (!planesReady_)
	{
		computePlanes();
		GA_HIP(hipStreamSynchronize(stream_));
	}
	return readPlanes_.as<uint64_t>();
}

void Context::uploadReads(const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens)
{
	GA_HIP(hipSetDevice(device_));
	const size_t n = seqs.size();
	ReadSet& rs = reads_;
	rs.offset.resize(n);
	rs.length.assign(lens.begin(), lens.end());
	rs.bytes = 0;
	rs.maxLen = 0;
	for (size_t i = 0; i < n; i++)
	{
		rs.offset[i] = rs.bytes;
		rs.bytes += lens[i];
		rs.maxLen = std::max(rs.maxLen, lens[i]);
	}
	rs.host.assign(rs.bytes + 1, 0);
	for (size_t i = 0; i < n; i++) std::memcpy(rs.host.data() + rs.offset[i], seqs[i], lens[i]);
	revBase_ = rs.bytes + 1;
	readBytes_.reserve(2 * (rs.bytes + 1));
	planesReady_ = false;
	revLo_ = revHi_ = 0;
	readOff_.reserve(std::max<size_t>(1, n) * sizeof(uint64_t));
	readLen_.reserve(std::max<size_t>(1, n) * sizeof(uint32_t));
	GA_HIP(hipMemcpyAsync(readBytes_.ptr, rs.host.data(), rs.bytes + 1, hipMemcpyHostToDevice, stream_));
	if (n > 0)
	{
		GA_HIP(hipMemcpyAsync(readOff_.ptr, rs.offset.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
		GA_HIP(hipMemcpyAsync(readLen_.ptr, rs.length.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
	}
	GA_HIP(hipStreamSynchronize(stream_));
}

void Context::writeReverse(uint32_t first, uint32_t count)
{
	// reverse complements (and the planes of the buffer holding them) once per
	// uploaded read range
	const bool revNeeded = !(first >= revLo_ && (uint64_t)first + count <= revHi_);
	if (!revNeeded || count == 0) return;
	const uint64_t* off = readOff_.as<uint64_t>() + first;
	const uint32_t* len = readLen_.as<uint32_t>() + first;
	GA_HIP(gad::launchRevcomp(stream_, readBytes_.as<char>(), readBytes_.as<char>() + revBase_, off, len, count, 0));
	// the reverse complements just written change the buffer's planes
	std::lock_guard<std::mutex> lock(cacheMutex_);
	computePlanes();
	if (revHi_ > revLo_ && first <= revHi_ && (uint64_t)first + count >= revLo_)
	{
		revLo_ = std::min<uint64_t>(revLo_, first);
		revHi_ = std::max<uint64_t>(revHi_, (uint64_t)first + count);
	}
	else
	{
		revLo_ = first;
		revHi_ = (uint64_t)first + count;
	}
}

void Context::beginAlignmentCall()
{
	std::lock_guard<std::mutex> lock(cacheMutex_);
	revLo_ = revHi_ = 0;
	planesReady_ = false;
}

void Context::prepareReads(uint32_t first, uint32_t count)
{
	GA_HIP(hipSetDevice(device_));
	if ((uint64_t)first + count > reads_.size()) throw std::runtime_error("read range outside the resident reads");
	writeReverse(first, count);
	GA_HIP(hipStreamSynchronize(stream_));
}

void Context::lookupSeeds(const MinimizerIndex& idx, uint32_t first, uint32_t count, SeedMatches& out)
{
	GA_HIP(hipSetDevice(device_));
	if (idx.table.empty()) throw std::runtime_error("minimizer index is empty");
	const ReadSet& rs = reads_;
	if ((uint64_t)first + count > rs.size()) throw std::runtime_error("read range outside the resident reads");
	const DeviceIndex& di = deviceIndex(idx);
	out.matchCount.assign(count, 0);
	out.matchOffset.assign(count, 0);
	out.matches.clear();
	out.staged = nullptr;
	out.lookupMs = 0;
	lookupFirst_ = first;
	lookupCount_ = count;
	if (count == 0) return;
	// matches of read i are written at its byte offset (at most one per base)
	matches_.reserve((rs.bytes + 1) * sizeof(gad::SeedMatch));
	matchCount_.reserve(count * sizeof(uint32_t));
	const uint64_t* off = readOff_.as<uint64_t>() + first;
	const uint32_t* len = readLen_.as<uint32_t>() + first;
	GA_HIP(hipEventRecord(ev0_, stream_));
	GA_HIP(gad::launchSeedLookup(stream_, readBytes_.as<char>(), off, len, count, di.table.as<gad::SeedTableSlot>(), di.mask, idx.k, idx.w, idx.maxCount, matches_.as<gad::SeedMatch>(), matchCount_.as<uint32_t>()));
	writeReverse(first, count);
	GA_HIP(hipEventRecord(ev1_, stream_));
	GA_HIP(hipMemcpyAsync(out.matchCount.data(), matchCount_.ptr, count * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	float ms = 0;
	GA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
	out.lookupMs = ms;
	// gather the used prefixes into one contiguous block and copy that
	uint64_t total = 0;
	for (uint32_t i = 0; i < count; i++)
	{
		if (out.matchCount[i] > rs.length[first + i]) throw HipError("seed lookup kernel returned an inconsistent match count");
		out.matchOffset[i] = total;
		total += out.matchCount[i];
	}
	if (total == 0) return;
	gathered_.reserve(total * sizeof(gad::SeedMatch));
	matchOff_.reserve(count * sizeof(uint64_t));
	GA_HIP(hipMemcpyAsync(matchOff_.ptr, out.matchOffset.data(), count * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
	GA_HIP(gad::launchGatherMatches(stream_, matches_.as<gad::SeedMatch>(), off, matchCount_.as<uint32_t>(), matchOff_.as<uint64_t>(), count, gathered_.as<gad::SeedMatch>()));
	matchesHost_.reserve(total * sizeof(gad::SeedMatch));
	GA_HIP(hipMemcpyAsync(matchesHost_.ptr, gathered_.ptr, total * sizeof(gad::SeedMatch), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	out.staged = matchesHost_.as<gad::SeedMatch>();
}

void Context::orderedSeeds(const GraphHost& g, const MinimizerIndex& idx, uint32_t first, uint32_t count, const SeedMatches& sm, double density,
                           std::vector<std::vector<SeedHit>>& seeds, std::vector<uint8_t>& asserted)
{
	GA_HIP(hipSetDevice(device_));
	if (first != lookupFirst_ || count != lookupCount_ || sm.matchCount.size() != count)
		throw std::runtime_error("orderedSeeds needs the matches of the last lookupSeeds");
	seeds.assign(count, {});
	asserted.assign(count, 0);
	if (count == 0) return;
	const DeviceGraph& dgb = deviceGraphBuffers(g);
	const DeviceIndex& di = deviceIndex(idx);
	// room per read: every position of every match (the density cut keeps fewer)
	std::vector<uint64_t> off(count), cap(count);
	uint64_t total = 0;
	const gad::SeedMatch* m = sm.data();
	for (uint32_t i = 0; i < count; i++)
	{
		uint64_t c = 0;
		for (uint32_t k = 0; k < sm.matchCount[i]; k++) c += m[sm.matchOffset[i] + k].count;
		off[i] = total;
		cap[i] = c;
		total += c;
	}
	seedOut_.reserve(std::max<uint64_t>(total, 1) * sizeof(gad::DevSeedHit));
	seedScratch_.reserve(std::max<uint64_t>(total, 1) * sizeof(gad::ChainKey));
	seedOff_.reserve(count * sizeof(uint64_t));
	seedCap_.reserve(count * sizeof(uint64_t));
	seedCount_.reserve(count * sizeof(int64_t));
	GA_HIP(hipMemcpyAsync(seedOff_.ptr, off.data(), count * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
	GA_HIP(hipMemcpyAsync(seedCap_.ptr, cap.data(), count * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
	// the lookup left the reads' matches gathered in gathered_ at matchOff_; the
	// kernel sorts each read's matches in place (the host copy is not touched)
	uint64_t nmatches = 0;
	for (uint32_t i = 0; i < count; i++) nmatches += sm.matchCount[i];
	// no matches: no read has seeds (lookupSeeds then left no gathered matches)
	if (nmatches == 0) return;
	if (gathered_.bytes < nmatches * sizeof(gad::SeedMatch) || matchOff_.bytes < count * sizeof(uint64_t)) throw std::runtime_error("gathered matches missing");
	const uint32_t* len = readLen_.as<uint32_t>() + first;
	GA_HIP(gad::launchSeedOrder(stream_, dgb.seedView, di.positions.as<uint64_t>(), di.npositions, idx.maxCount, idx.k, gathered_.as<gad::SeedMatch>(),
	                            matchOff_.as<uint64_t>(), matchCount_.as<uint32_t>(), len, count, density, seedOut_.as<gad::DevSeedHit>(),
	                            seedScratch_.as<gad::ChainKey>(), seedOff_.as<uint64_t>(), seedCap_.as<uint64_t>(), seedCount_.as<int64_t>()));
	std::vector<int64_t> n(count);
	std::vector<gad::DevSeedHit> hits(total);
	GA_HIP(hipMemcpyAsync(n.data(), seedCount_.ptr, count * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
	if (total > 0) GA_HIP(hipMemcpyAsync(hits.data(), seedOut_.ptr, total * sizeof(gad::DevSeedHit), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	for (uint32_t i = 0; i < count; i++)
	{
		if (n[i] == -1)
		{
			asserted[i] = 1;
			continue;
		}
		if (n[i] < 0 || (uint64_t)n[i] > cap[i]) throw HipError("seed ordering kernel returned an inconsistent seed count");
		auto& v = seeds[i];
		v.resize(n[i]);
		for (int64_t k = 0; k < n[i]; k++)
		{
			const gad::DevSeedHit& d = hits[off[i] + k];
			SeedHit& h = v[k];
			h.nodeID = d.nodeID;
			h.nodeOffset = d.nodeOffset;
			h.seqPos = d.seqPos;
			h.matchLen = d.matchLen;
			h.reverse = d.reverse != 0;
			h.alignmentGraphNodeId = d.alignmentGraphNodeId;
			h.alignmentGraphNodeOffset = d.alignmentGraphNodeOffset;
			h.rawSeedGoodness = d.rawSeedGoodness;
			h.seedGoodness = d.seedGoodness;
			h.seedClusterSize = d.seedClusterSize;
		}
	}
}

void Context::runOnce(ExtendLane& L, const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out, const char* devSeqBase, const uint64_t* devPlanes)
{
	// split into launches whose workspaces fit the memory budget
	size_t begin = 0;
	while (begin < which.size())
	{
		uint64_t bytes = 0;
		size_t end = begin;
		while (end < which.size())
		{
			uint64_t b = gad::requestArenaBytes(jobs[which[end]].len, caps);
			if (end > begin && bytes + b > arenaBudgetBytes / lanes_.size()) break;
			bytes += b;
			end++;
		}
		std::vector<uint32_t> part(which.begin() + begin, which.begin() + end);
		runLaunch(L, dg, p, jobs, part, caps, wantSummaries, out, devSeqBase, devPlanes);
		begin = end;
	}
}

void Context::runLaunch(ExtendLane& L, const gad::DevGraph& dg, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, const std::vector<uint32_t>& which, const gad::ArenaCaps& caps, bool wantSummaries, ExtBatch& out, const char* devSeqBase, const uint64_t* devPlanes)
{
	const uint32_t n = which.size();
	if (n == 0) return;
	std::vector<gad::ExtRequest> reqs(n);
	std::vector<uint64_t> arenaOff(n), sumOff(n);
	std::vector<int32_t> bandHost;
	uint64_t seqBytes = 0, traceItems = 0, sumItems = 0, arenaTotal = 0;
	for (uint32_t k = 0; k < n; k++)
	{
		const ExtJob& j = jobs[which[k]];
		reqs[k].seqOffset = devSeqBase ? j.devOffset : seqBytes;
		reqs[k].traceOffset = traceItems;
		reqs[k].len = j.len;
		reqs[k].startNode = j.startNode;
		reqs[k].startOffset = j.startOffset;
		reqs[k].traceCap = traceCapFor(j.len);
		if (j.bandLen > 0)
		{
			// a schedule longer than the extension's slices is cut to them
			const uint32_t used = std::min<uint32_t>(j.bandLen, j.len / 64 + 1);
			reqs[k].bandOff = (uint32_t)bandHost.size();
			reqs[k].bandLen = used;
			bandHost.insert(bandHost.end(), j.bands, j.bands + used);
		}
		arenaOff[k] = arenaTotal;
		sumOff[k] = sumItems;
		seqBytes += j.len;
		traceItems += traceCapFor(j.len);
		sumItems += j.len / 64 + 2;
		arenaTotal += gad::requestArenaBytes(j.len, caps);
	}
	std::vector<char> seqHost;
	if (!devSeqBase)
	{
		seqHost.resize(seqBytes + 1);
		for (uint32_t k = 0; k < n; k++) std::memcpy(seqHost.data() + reqs[k].seqOffset, jobs[which[k]].seq, reqs[k].len);
	}
	L.arena.reserve(arenaTotal);
	L.arenaOffsets.reserve(n * sizeof(uint64_t));
	L.reqs.reserve(n * sizeof(gad::ExtRequest));
	if (!devSeqBase) L.seqs.reserve(seqHost.size());
	L.traces.reserve(traceItems * sizeof(uint64_t));
	L.results.reserve(n * sizeof(gad::ExtResult));
	if (wantSummaries)
	{
		L.sums.reserve(sumItems * sizeof(gad::SliceSummary));
		L.sumOffsets.reserve(n * sizeof(uint64_t));
		GA_HIP(hipMemcpyAsync(L.sumOffsets.ptr, sumOff.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, L.stream));
	}
	GA_HIP(hipMemcpyAsync(L.arenaOffsets.ptr, arenaOff.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, L.stream));
	GA_HIP(hipMemcpyAsync(L.reqs.ptr, reqs.data(), n * sizeof(gad::ExtRequest), hipMemcpyHostToDevice, L.stream));
	gad::ExtParams lp = p;
	if (!bandHost.empty())
	{
		L.bands.reserve(bandHost.size() * sizeof(int32_t));
		GA_HIP(hipMemcpyAsync(L.bands.ptr, bandHost.data(), bandHost.size() * sizeof(int32_t), hipMemcpyHostToDevice, L.stream));
		lp.bandSched = L.bands.as<int32_t>();
	}
	const uint64_t* planes = devPlanes;
	if (!devSeqBase)
	{
		GA_HIP(hipMemcpyAsync(L.seqs.ptr, seqHost.data(), seqHost.size(), hipMemcpyHostToDevice, L.stream));
		L.planes.reserve(gad::eqPlaneWords(seqHost.size()) * sizeof(uint64_t));
		GA_HIP(gad::launchEqPlanes(L.stream, L.seqs.as<char>(), seqHost.size(), L.planes.as<uint64_t>()));
		planes = L.planes.as<uint64_t>();
	}
	if (!planes) throw std::runtime_error("extension launch without match-mask planes");
	GA_HIP(hipEventRecord(L.ev0, L.stream));
	GA_HIP(gad::launchExtend(L.stream, n, lanesPerWave(n), kernelWavesPerSimd, dg, lp, L.reqs.as<gad::ExtRequest>(), devSeqBase ? devSeqBase : L.seqs.as<char>(), L.traces.as<uint64_t>(), L.results.as<gad::ExtResult>(), L.arena.as<unsigned char>(), L.arenaOffsets.as<uint64_t>(), caps, wantSummaries ? L.sums.as<gad::SliceSummary>() : nullptr, wantSummaries ? L.sumOffsets.as<uint64_t>() : nullptr, planes, L.capWaves));
	GA_HIP(hipEventRecord(L.ev1, L.stream));
	std::vector<gad::ExtResult> res(n);
	GA_HIP(hipMemcpyAsync(res.data(), L.results.ptr, n * sizeof(gad::ExtResult), hipMemcpyDeviceToHost, L.stream));
	std::vector<gad::SliceSummary> sm;
	if (wantSummaries)
	{
		sm.resize(sumItems);
		GA_HIP(hipMemcpyAsync(sm.data(), L.sums.ptr, sumItems * sizeof(gad::SliceSummary), hipMemcpyDeviceToHost, L.stream));
	}
	GA_HIP(hipStreamSynchronize(L.stream));
	float ms = 0;
	GA_HIP(hipEventElapsedTime(&ms, L.ev0, L.ev1));
	out.kernelMs += ms;
	out.launches++;
	// traces: encode the used prefix of every kept job's trace region into one
	// contiguous block of code bytes on the device, then copy only that block
	// and the escaped items
	std::vector<uint64_t> srcOff, dstOff;
	std::vector<uint32_t> cnt;
	const uint64_t base = L.hostCodesUsed;
	uint64_t total = 0;
	for (uint32_t k = 0; k < n; k++)
	{
		uint32_t jid = which[k];
		const gad::ExtResult& r = res[k];
		if (r.status < gad::EXT_OK || r.status > gad::EXT_OVERFLOW || r.ntrace > reqs[k].traceCap || r.nslices > reqs[k].len / 64 + 1)
			throw HipError("extension kernel returned an inconsistent result record");
		out.results[jid] = r;
		if (r.status == gad::EXT_OVERFLOW) continue;
		out.wordSteps += r.wordSteps;
		out.traceOffset[jid] = base + total;
		if (r.ntrace > 0)
		{
			srcOff.push_back(reqs[k].traceOffset);
			dstOff.push_back(total);
			cnt.push_back(r.ntrace);
			total += r.ntrace;
		}
		if (wantSummaries)
		{
			out.summaryOffset[jid] = out.summaries.size();
			out.summaries.insert(out.summaries.end(), sm.begin() + sumOff[k], sm.begin() + sumOff[k] + r.nslices);
		}
	}
	if (base + total > L.hostCodes.bytes) throw HipError("trace codes exceed the host staging buffer");
	L.hostCodesUsed = base + total;
	if (total > 0)
	{
		const uint32_t m = cnt.size();
		L.gatherSrc.reserve(m * sizeof(uint64_t));
		L.gatherDst.reserve(m * sizeof(uint64_t));
		L.gatherCnt.reserve(m * sizeof(uint32_t));
		L.traceCodes.reserve(total);
		L.traceEscapes.reserve(total * 2 * sizeof(uint64_t));
		L.escapeCount.reserve(sizeof(uint32_t));
		GA_HIP(hipMemcpyAsync(L.gatherSrc.ptr, srcOff.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, L.stream));
		GA_HIP(hipMemcpyAsync(L.gatherDst.ptr, dstOff.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, L.stream));
		GA_HIP(hipMemcpyAsync(L.gatherCnt.ptr, cnt.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, L.stream));
		GA_HIP(hipMemsetAsync(L.escapeCount.ptr, 0, sizeof(uint32_t), L.stream));
		GA_HIP(gad::launchTraceCodes(L.stream, dg, L.traces.as<uint64_t>(), L.gatherSrc.as<uint64_t>(), L.gatherCnt.as<uint32_t>(), L.gatherDst.as<uint64_t>(), m, L.traceCodes.as<uint8_t>(), L.traceEscapes.as<uint64_t>(), L.escapeCount.as<uint32_t>(), total));
		uint32_t nesc = 0;
		GA_HIP(hipMemcpyAsync(&nesc, L.escapeCount.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, L.stream));
		GA_HIP(hipMemcpyAsync(L.hostCodes.as<uint8_t>() + base, L.traceCodes.ptr, total, hipMemcpyDeviceToHost, L.stream));
		GA_HIP(hipStreamSynchronize(L.stream));
		if (nesc > total) throw HipError("trace encoder returned an inconsistent escape count");
		std::vector<uint64_t> esc(2 * (size_t)nesc);
		if (nesc > 0)
		{
			GA_HIP(hipMemcpyAsync(esc.data(), L.traceEscapes.ptr, esc.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, L.stream));
			GA_HIP(hipStreamSynchronize(L.stream));
		}
		for (uint32_t k = 0; k < nesc; k++) out.escapes.emplace_back(base + esc[2 * k], esc[2 * k + 1]);
	}
}

void Context::extend(const GraphHost& g, const gad::ExtParams& p, const std::vector<ExtJob>& jobs, bool wantSummaries, ExtBatch& out, const char* devSeqBase, uint32_t lane)
{
	if (lane >= lanes_.size()) throw std::runtime_error("extension lane out of range");
	ExtendLane& L = *lanes_[lane];
	GA_HIP(hipSetDevice(device_));
	const gad::DevGraph& dg = deviceGraph(g);
	const uint64_t* devPlanes = devSeqBase ? residentPlanes(devSeqBase) : nullptr;
	out.results.assign(jobs.size(), gad::ExtResult{});
	out.traceOffset.assign(jobs.size(), 0);
	out.summaryOffset.assign(jobs.size(), 0);
	out.escapes.clear();
	out.summaries.clear();
	// every job's codes land once (overflowed attempts copy none), so the sum of
	// the trace capacities bounds this call's code bytes
	uint64_t codeBound = 0;
	for (const ExtJob& j : jobs) codeBound += traceCapFor(j.len);
	L.hostCodes.reserve(codeBound);
	L.hostCodesUsed = 0;
	out.codes = L.hostCodes.as<uint8_t>();
	std::vector<uint32_t> which(jobs.size());
	for (uint32_t i = 0; i < jobs.size(); i++) which[i] = i;
	// longest first: the 64 threads of a wave then carry similar lengths
	std::stable_sort(which.begin(), which.end(), [&](uint32_t a, uint32_t b) { return jobs[a].len > jobs[b].len; });
	gad::ArenaCaps caps = initialArenaCaps();
	for (int attempt = 0; attempt < kArenaAttempts && !which.empty(); attempt++)
	{
		runOnce(L, dg, p, jobs, which, caps, wantSummaries, out, devSeqBase, devPlanes);
		std::vector<uint32_t> again;
		for (uint32_t id : which)
			if (out.results[id].status == gad::EXT_OVERFLOW) again.push_back(id);
		which.swap(again);
		if (attempt + 1 < kArenaAttempts) out.reruns += which.size();
		growArenaCaps(caps);
	}
	std::sort(out.escapes.begin(), out.escapes.end());
}

static uint64_t roundUp256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

const gad::DevGraph* Context::residentGraph(const GraphHost& g)
{
	std::lock_guard<std::mutex> lock(cacheMutex_);
	for (auto& c : graphs_)
		if (c->uid == g.uid) return &c->view;
	return nullptr;
}

namespace {

// Host staging of the no-allocation entry point: fixed arrays, one call at a
// time per process (calls on one GPU share its device scratch anyway).
struct NoAllocStaging
{
	std::mutex m;
	gad::ExtRequest reqs[Context::kNoAllocMaxRequests];
	uint64_t arenaOff[Context::kNoAllocMaxRequests];
	gad::ExtResult res[Context::kNoAllocMaxRequests];
	uint32_t job[Context::kNoAllocMaxRequests];
};
NoAllocStaging gNoAlloc;

} // namespace

bool Context::extendNoAlloc(const gad::DevGraph& dg, const gad::ExtParams& p, uint32_t n, const char* const* seqs, const uint32_t* lens, const uint32_t* startNodes, const uint32_t* startOffsets, int64_t traceCap, int64_t* traces, gad::ExtResult* results, double* kernelMs)
{
	if (n > kNoAllocMaxRequests || !scratch_) return false;
	std::lock_guard<std::mutex> lock(gNoAlloc.m);
	NoAllocStaging& S = gNoAlloc;
	// scratch layout: requests | workspace offsets | results | sequences |
	// planes | traces | workspaces, every part on a 256-byte boundary
	uint64_t seqBytes = 0, traceItems = 0;
	for (uint32_t i = 0; i < n; i++)
	{
		seqBytes += lens[i];
		traceItems += traceCapFor(lens[i]);
	}
	const uint64_t offReqs = 0;
	const uint64_t offArenaOff = offReqs + roundUp256((uint64_t)n * sizeof(gad::ExtRequest));
	const uint64_t offResults = offArenaOff + roundUp256((uint64_t)n * sizeof(uint64_t));
	const uint64_t offSeqs = offResults + roundUp256((uint64_t)n * sizeof(gad::ExtResult));
	const uint64_t offPlanes = offSeqs + roundUp256(seqBytes + 1);
	const uint64_t offTraces = offPlanes + roundUp256(gad::eqPlaneWords(seqBytes + 1) * sizeof(uint64_t));
	const uint64_t offArenas = offTraces + roundUp256(traceItems * sizeof(uint64_t));
	gad::ArenaCaps caps = initialArenaCaps();
	uint64_t arenaTotal = 0;
	for (uint32_t i = 0; i < n; i++) arenaTotal += roundUp256(gad::requestArenaBytes(lens[i], caps));
	if (offArenas + arenaTotal > gad::kNoAllocScratchBytes) return false;
	GA_HIP(hipSetDevice(device_));
	unsigned char* base = scratch_;
	char* dSeqs = (char*)(base + offSeqs);
	uint64_t* dPlanes = (uint64_t*)(base + offPlanes);
	uint64_t* dTraces = (uint64_t*)(base + offTraces);
	uint64_t at = 0, tat = 0;
	for (uint32_t i = 0; i < n; i++)
	{
		if (lens[i] > 0) GA_HIP(hipMemcpyAsync(dSeqs + at, seqs[i], lens[i], hipMemcpyHostToDevice, stream_));
		gad::ExtRequest& r = S.reqs[i];
		r.seqOffset = at;
		r.traceOffset = tat;
		r.len = lens[i];
		r.startNode = startNodes[i];
		r.startOffset = startOffsets[i];
		r.traceCap = traceCapFor(lens[i]);
		S.job[i] = i;
		at += lens[i];
		tat += r.traceCap;
	}
	// the planes of the last block read the byte after the last sequence: zero it
	static const char zero = 0;
	GA_HIP(hipMemcpyAsync(dSeqs + at, &zero, 1, hipMemcpyHostToDevice, stream_));
	GA_HIP(gad::launchEqPlanes(stream_, dSeqs, seqBytes + 1, dPlanes));
	float msTotal = 0;
	// first attempt over every request, then reruns of the overflowed ones with
	// larger workspaces, each as long as the scratch holds their workspaces
	uint32_t m = n;
	for (int attempt = 0; attempt < kArenaAttempts && m > 0; attempt++)
	{
		uint64_t aoff = 0;
		for (uint32_t k = 0; k < m; k++)
		{
			S.arenaOff[k] = aoff;
			aoff += roundUp256(gad::requestArenaBytes(S.reqs[k].len, caps));
		}
		if (offArenas + aoff > gad::kNoAllocScratchBytes) break;
		GA_HIP(hipMemcpyAsync(base + offReqs, S.reqs, (size_t)m * sizeof(gad::ExtRequest), hipMemcpyHostToDevice, stream_));
		GA_HIP(hipMemcpyAsync(base + offArenaOff, S.arenaOff, (size_t)m * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
		GA_HIP(hipEventRecord(ev0_, stream_));
		GA_HIP(gad::launchExtend(stream_, m, lanesPerWave(m), kernelWavesPerSimd, dg, p, (const gad::ExtRequest*)(base + offReqs), dSeqs, dTraces, (gad::ExtResult*)(base + offResults), base + offArenas, (const uint64_t*)(base + offArenaOff), caps, nullptr, nullptr, dPlanes));
		GA_HIP(hipEventRecord(ev1_, stream_));
		GA_HIP(hipMemcpyAsync(S.res, base + offResults, (size_t)m * sizeof(gad::ExtResult), hipMemcpyDeviceToHost, stream_));
		GA_HIP(hipStreamSynchronize(stream_));
		float ms = 0;
		GA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
		msTotal += ms;
		uint32_t again = 0;
		for (uint32_t k = 0; k < m; k++)
		{
			const gad::ExtResult& r = S.res[k];
			if (r.status < gad::EXT_OK || r.status > gad::EXT_OVERFLOW || r.ntrace > S.reqs[k].traceCap)
				throw HipError("extension kernel returned an inconsistent result record");
			const uint32_t j = S.job[k];
			results[j] = r;
			if (r.status == gad::EXT_OVERFLOW)
			{
				S.reqs[again] = S.reqs[k];
				S.job[again] = j;
				again++;
				continue;
			}
			// the packed items into the front of the caller's trace block, then
			// unpacked in place from the last item down (item k's four words
			// start at or after packed word k, which is read first)
			int64_t* t = traces + 4 * (int64_t)j * traceCap;
			if (r.ntrace > 0) GA_HIP(hipMemcpyAsync(t, dTraces + S.reqs[k].traceOffset, (size_t)r.ntrace * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
		}
		GA_HIP(hipStreamSynchronize(stream_));
		for (uint32_t k = 0; k < m; k++)
		{
			const gad::ExtResult& r = S.res[k];
			if (r.status == gad::EXT_OVERFLOW) continue;
			int64_t* t = traces + 4 * (int64_t)S.job[k] * traceCap;
			for (int64_t q = (int64_t)r.ntrace - 1; q >= 0; q--)
			{
				const uint64_t v = (uint64_t)t[q];
				const uint32_t hi = (uint32_t)(v >> 32);
				t[4 * q] = (uint32_t)v;
				t[4 * q + 1] = (hi >> 1) & 63;
				t[4 * q + 2] = (int64_t)(hi >> 7) - 1;
				t[4 * q + 3] = hi & 1;
			}
		}
		m = again;
		growArenaCaps(caps);
	}
	if (kernelMs) *kernelMs = msTotal;
	return true;
}

static uint64_t readLoopBytes(uint32_t len, const ReadLoopCaps& caps)
{
	return roundUp256(gad::requestArenaBytes(len, caps.arena)) + 2ull * traceCapFor(len) * 8 + (uint64_t)poolCapFor(len, caps) * 8 + (uint64_t)caps.alnCap * sizeof(gad::DevAlignment);
}

void Context::runReadLoops(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out)
{
	static_assert(sizeof(gad::DevAlignment) == 32, "alignment records are gathered as four 64-bit words");
	GA_HIP(hipSetDevice(device_));
	const gad::DevGraph& dg = deviceGraph(g);
	const size_t n = jobs.size();
	out = ReadLoopOutput{};
	out.results.assign(n, gad::ReadResult{});
	out.alnOffset.assign(n, 0);
	out.poolOffset.assign(n, 0);
	if (n == 0) return;
	for (const auto& j : jobs)
	{
		if (j.read >= reads_.size()) throw std::runtime_error("read loop job names a read that is not resident");
		if (j.seedStart + j.seedCount > seeds.size()) throw std::runtime_error("read loop job seeds out of range");
	}
	rlSeeds_.reserve(std::max<size_t>(1, seeds.size()) * sizeof(gad::DevSeed));
	if (!seeds.empty()) GA_HIP(hipMemcpyAsync(rlSeeds_.ptr, seeds.data(), seeds.size() * sizeof(gad::DevSeed), hipMemcpyHostToDevice, stream_));
	std::vector<uint32_t> which(n);
	for (uint32_t i = 0; i < n; i++) which[i] = i;
	// longest reads first: the dispatcher starts the longest seed loops earliest
	std::stable_sort(which.begin(), which.end(), [&](uint32_t a, uint32_t b) { return reads_.length[jobs[a].read] > reads_.length[jobs[b].read]; });
	ReadLoopCaps caps = initialReadLoopCaps();
	for (int attempt = 0; attempt < kArenaAttempts && !which.empty(); attempt++)
	{
		size_t begin = 0;
		while (begin < which.size())
		{
			uint64_t bytes = 0;
			size_t end = begin;
			while (end < which.size())
			{
				uint64_t b = readLoopBytes(reads_.length[jobs[which[end]].read], caps);
				if (end > begin && bytes + b > arenaBudgetBytes) break;
				bytes += b;
				end++;
			}
			std::vector<uint32_t> part(which.begin() + begin, which.begin() + end);
			runReadLoopLaunch(dg, p, jobs, part, caps, out);
			begin = end;
		}
		std::vector<uint32_t> again;
		for (uint32_t id : which)
			if (out.results[id].status == gad::READLOOP_OVERFLOW) again.push_back(id);
		which.swap(again);
		if (attempt + 1 < kArenaAttempts) out.reruns += which.size();
		growReadLoopCaps(caps);
	}
}

void Context::runReadLoopLaunch(const gad::DevGraph& dg, const gad::ReadLoopParams& p, const std::vector<ReadLoopJob>& jobs, const std::vector<uint32_t>& which, const ReadLoopCaps& caps, ReadLoopOutput& out)
{
	const uint32_t n = which.size();
	if (n == 0) return;
	std::vector<gad::ReadJob> rj(n);
	uint64_t arenaTotal = 0, traceTotal = 0, poolTotal = 0, alnTotal = 0;
	for (uint32_t k = 0; k < n; k++)
	{
		const ReadLoopJob& j = jobs[which[k]];
		const uint32_t len = reads_.length[j.read];
		gad::ReadJob& r = rj[k];
		r.fwdOffset = reads_.offset[j.read];
		r.revOffset = revBase_ + reads_.offset[j.read];
		r.seedStart = j.seedStart;
		r.arenaOffset = arenaTotal;
		r.traceOffset = traceTotal;
		r.poolOffset = poolTotal;
		r.alnOffset = alnTotal;
		r.extendSeeds = j.extendSeeds;
		r.len = len;
		r.seedCount = j.seedCount;
		r.traceCap = traceCapFor(len);
		r.poolCap = poolCapFor(len, caps);
		r.alnCap = caps.alnCap;
		r.pad = 0;
		arenaTotal += roundUp256(gad::requestArenaBytes(len, caps.arena));
		traceTotal += 2ull * r.traceCap;
		poolTotal += r.poolCap;
		alnTotal += r.alnCap;
	}
	arena_.reserve(arenaTotal);
	rlJobs_.reserve(n * sizeof(gad::ReadJob));
	rlTraces_.reserve(traceTotal * 8);
	rlPool_.reserve(std::max<uint64_t>(1, poolTotal) * 8);
	rlAlns_.reserve(alnTotal * sizeof(gad::DevAlignment));
	rlResults_.reserve(n * sizeof(gad::ReadResult));
	GA_HIP(hipMemcpyAsync(rlJobs_.ptr, rj.data(), n * sizeof(gad::ReadJob), hipMemcpyHostToDevice, stream_));
	GA_HIP(hipEventRecord(ev0_, stream_));
	GA_HIP(gad::launchReadLoops(stream_, n, lanesPerWave(n), kernelWavesPerSimd, dg, p, rlJobs_.as<gad::ReadJob>(), readBytes_.as<char>(), rlSeeds_.as<gad::DevSeed>(), rlTraces_.as<uint64_t>(), rlPool_.as<uint64_t>(), rlAlns_.as<gad::DevAlignment>(), rlResults_.as<gad::ReadResult>(), arena_.as<unsigned char>(), caps.arena, residentPlanes(readBytes_.as<char>())));
	GA_HIP(hipEventRecord(ev1_, stream_));
	std::vector<gad::ReadResult> res(n);
	GA_HIP(hipMemcpyAsync(res.data(), rlResults_.ptr, n * sizeof(gad::ReadResult), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	float ms = 0;
	GA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
	out.kernelMs += ms;
	out.launches++;
	collectReadLoops(rj, res, which, out, rlAlns_.as<uint64_t>(), rlPool_.as<uint64_t>());
}


void Context::collectReadLoops(const std::vector<gad::ReadJob>& rj, const std::vector<gad::ReadResult>& res, const std::vector<uint32_t>& which, ReadLoopOutput& out, const uint64_t* alnSrc, const uint64_t* poolSrc)
{
	const uint32_t n = which.size();
	// gather every finished read's alignment records and used pool prefix
	std::vector<uint64_t> srcOff, dstOff;
	std::vector<uint32_t> cnt;
	std::vector<uint64_t> psrc, pdst;
	std::vector<uint32_t> pcnt;
	const uint64_t alnBase = out.alns.size(), poolBase = out.pool.size();
	uint64_t alnUsed = 0, poolUsed = 0;
	for (uint32_t k = 0; k < n; k++)
	{
		const gad::ReadResult& r = res[k];
		const uint32_t jid = which[k];
		if ((r.status != gad::READLOOP_DONE && r.status != gad::READLOOP_ASSERT && r.status != gad::READLOOP_OVERFLOW) || r.nAlignments > rj[k].alnCap || r.poolUsed > rj[k].poolCap)
			throw HipError("read loop kernel returned an inconsistent result record");
		out.results[jid] = r;
		out.wordSteps += r.wordSteps;
		out.extensions += r.extensions;
		if (r.status != gad::READLOOP_DONE) continue;
		out.alnOffset[jid] = alnBase + alnUsed;
		out.poolOffset[jid] = poolBase + poolUsed;
		if (r.nAlignments > 0)
		{
			srcOff.push_back(rj[k].alnOffset * 4);
			dstOff.push_back(alnUsed * 4);
			cnt.push_back(r.nAlignments * 4);
			alnUsed += r.nAlignments;
		}
		if (r.poolUsed > 0)
		{
			psrc.push_back(rj[k].poolOffset);
			pdst.push_back(poolUsed);
			pcnt.push_back(r.poolUsed);
			poolUsed += r.poolUsed;
		}
	}
	out.alns.resize(alnBase + alnUsed);
	out.pool.resize(poolBase + poolUsed);
	auto gather = [&](const uint64_t* src, const std::vector<uint64_t>& so, const std::vector<uint64_t>& dof, const std::vector<uint32_t>& c, uint64_t total, void* dst) {
		if (total == 0) return;
		const uint32_t m = c.size();
		gatherSrc_.reserve(m * sizeof(uint64_t));
		gatherDst_.reserve(m * sizeof(uint64_t));
		gatherCnt_.reserve(m * sizeof(uint32_t));
		traceOut_.reserve(total * sizeof(uint64_t));
		GA_HIP(hipMemcpyAsync(gatherSrc_.ptr, so.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
		GA_HIP(hipMemcpyAsync(gatherDst_.ptr, dof.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
		GA_HIP(hipMemcpyAsync(gatherCnt_.ptr, c.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
		GA_HIP(gad::launchGatherU64(stream_, src, gatherSrc_.as<uint64_t>(), gatherCnt_.as<uint32_t>(), gatherDst_.as<uint64_t>(), m, traceOut_.as<uint64_t>()));
		GA_HIP(hipMemcpyAsync(dst, traceOut_.ptr, total * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
		GA_HIP(hipStreamSynchronize(stream_));
	};
	gather(alnSrc, srcOff, dstOff, cnt, alnUsed * 4, out.alns.data() + alnBase);
	gather(poolSrc, psrc, pdst, pcnt, poolUsed, out.pool.data() + poolBase);
}


float Context::runPersistentBatch(const gad::DevGraph& dg, const gad::ReadLoopParams& p, const std::vector<gad::ReadJob>& rj, uint32_t workers, const ReadLoopCaps& caps, uint32_t maxLen, std::vector<gad::ReadResult>& res)
{
	const uint32_t n = rj.size();
	uint64_t poolTotal = 0, alnTotal = 0;
	for (const auto& j : rj)
	{
		poolTotal = std::max<uint64_t>(poolTotal, j.poolOffset + j.poolCap);
		alnTotal = std::max<uint64_t>(alnTotal, j.alnOffset + j.alnCap);
	}
	const uint64_t arenaStride = roundUp256(gad::requestArenaBytes(maxLen, caps.arena));
	const uint64_t rawStride = 2ull * traceCapFor(maxLen);
	bJobs_.reserve(std::max<size_t>(1, n) * sizeof(gad::ReadJob));
	bCounter_.reserve(2 * sizeof(uint32_t));
	bTraces_.reserve((uint64_t)workers * rawStride * 8);
	bArena_.reserve((uint64_t)workers * arenaStride);
	bPool_.reserve(std::max<uint64_t>(1, poolTotal) * 8);
	bAlns_.reserve(std::max<uint64_t>(1, alnTotal) * sizeof(gad::DevAlignment));
	bResults_.reserve(std::max<size_t>(1, n) * sizeof(gad::ReadResult));
	const uint64_t* planes = residentPlanes(readBytes_.as<char>());
	uint32_t* counter = bCounter_.as<uint32_t>();
	// the launch shape: buffers, strides, workers and parameters; a graph captured
	// for one shape is replayed for every batch of that shape (the job count is
	// read from device memory)
	PersistentShape shape{};
	const void* ptrs[9] = {bJobs_.ptr, bCounter_.ptr, bTraces_.ptr, bArena_.ptr, bPool_.ptr, bAlns_.ptr, bResults_.ptr, readBytes_.ptr, rlSeeds_.ptr};
	std::memcpy(shape.ptrs, ptrs, sizeof(ptrs));
	shape.workers = workers;
	shape.rawStride = rawStride;
	shape.arenaStride = arenaStride;
	shape.caps = caps.arena;
	shape.p = p;
	shape.dg = dg;
	(void)planes;
	if (!graphExec_ || std::memcmp(&shape, &graphShape_, sizeof(shape)) != 0)
	{
		if (graphExec_) GA_HIP(hipGraphExecDestroy(graphExec_));
		if (graph_) GA_HIP(hipGraphDestroy(graph_));
		graphExec_ = nullptr;
		graph_ = nullptr;
		GA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
		hipError_t le = gad::launchReadLoopsPersistent(stream_, workers, dg, p, bJobs_.as<gad::ReadJob>(), counter + 1, counter, readBytes_.as<char>(), rlSeeds_.as<gad::DevSeed>(), bTraces_.as<uint64_t>(), rawStride, bPool_.as<uint64_t>(), bAlns_.as<gad::DevAlignment>(), bResults_.as<gad::ReadResult>(), bArena_.as<unsigned char>(), arenaStride, caps.arena, planes);
		hipGraph_t gr = nullptr;
		hipError_t ce = hipStreamEndCapture(stream_, &gr);
		GA_HIP(le);
		GA_HIP(ce);
		graph_ = gr;
		GA_HIP(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
		graphShape_ = shape;
		graphCaptures++;
	}
	const uint32_t count = n;
	if (n > 0) GA_HIP(hipMemcpyAsync(bJobs_.ptr, rj.data(), n * sizeof(gad::ReadJob), hipMemcpyHostToDevice, stream_));
	GA_HIP(hipMemcpyAsync(counter + 1, &count, sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
	GA_HIP(hipEventRecord(ev0_, stream_));
	GA_HIP(hipGraphLaunch(graphExec_, stream_));
	GA_HIP(hipEventRecord(ev1_, stream_));
	graphLaunches++;
	res.assign(n, gad::ReadResult{});
	if (n > 0) GA_HIP(hipMemcpyAsync(res.data(), bResults_.ptr, n * sizeof(gad::ReadResult), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	float ms = 0;
	GA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
	return ms;
}

void Context::runReadLoopsBatched(const GraphHost& g, const gad::ReadLoopParams& p, const std::vector<gad::DevSeed>& seeds, const std::vector<ReadLoopJob>& jobs, ReadLoopOutput& out)
{
	static_assert(sizeof(gad::DevAlignment) == 32, "alignment records are gathered as four 64-bit words");
	GA_HIP(hipSetDevice(device_));
	const gad::DevGraph& dg = deviceGraph(g);
	const size_t n = jobs.size();
	out = ReadLoopOutput{};
	out.results.assign(n, gad::ReadResult{});
	out.alnOffset.assign(n, 0);
	out.poolOffset.assign(n, 0);
	if (n == 0) return;
	for (const auto& j : jobs)
	{
		if (j.read >= reads_.size()) throw std::runtime_error("read loop job names a read that is not resident");
		if (j.seedStart + j.seedCount > seeds.size()) throw std::runtime_error("read loop job seeds out of range");
	}
	rlSeeds_.reserve(std::max<size_t>(1, seeds.size()) * sizeof(gad::DevSeed));
	if (!seeds.empty()) GA_HIP(hipMemcpyAsync(rlSeeds_.ptr, seeds.data(), seeds.size() * sizeof(gad::DevSeed), hipMemcpyHostToDevice, stream_));
	// longest reads first: the workers take the longest seed loops earliest
	std::vector<uint32_t> which(n);
	for (uint32_t i = 0; i < n; i++) which[i] = i;
	std::stable_sort(which.begin(), which.end(), [&](uint32_t a, uint32_t b) { return reads_.length[jobs[a].read] > reads_.length[jobs[b].read]; });
	// first attempt: the band queue in LDS and per-extension workspaces as the
	// single-lane extension kernel has them; overflows go to runReadLoops
	ReadLoopCaps caps = initialReadLoopCaps();
	caps.arena = initialArenaCaps();
	auto makeJobs = [&](const std::vector<uint32_t>& ids, std::vector<gad::ReadJob>& rj) {
		rj.resize(ids.size());
		uint64_t poolTotal = 0, alnTotal = 0;
		uint32_t maxLen = 0;
		for (size_t k = 0; k < ids.size(); k++)
		{
			const ReadLoopJob& j = jobs[ids[k]];
			const uint32_t len = reads_.length[j.read];
			gad::ReadJob& r = rj[k];
			r.fwdOffset = reads_.offset[j.read];
			r.revOffset = revBase_ + reads_.offset[j.read];
			r.seedStart = j.seedStart;
			r.arenaOffset = 0;
			r.traceOffset = 0;
			r.poolOffset = poolTotal;
			r.alnOffset = alnTotal;
			r.extendSeeds = j.extendSeeds;
			r.len = len;
			r.seedCount = j.seedCount;
			r.traceCap = traceCapFor(len);
			r.poolCap = poolCapFor(len, caps);
			r.alnCap = caps.alnCap;
			r.pad = 0;
			poolTotal += r.poolCap;
			alnTotal += r.alnCap;
			maxLen = std::max(maxLen, len);
		}
		return maxLen;
	};
	// the worker count: candidates of 2..8 resident workers per SIMD, timed on a
	// sample of every k-th read of the first batch (their outputs discarded)
	if (batchedWorkers == 0)
	{
		auto t0 = std::chrono::steady_clock::now();
		uint32_t forced = 0;
		if (const char* v = std::getenv("GA_BATCH_WAVES_PER_SIMD"))
		{
			const int x = std::atoi(v);
			if (x >= 1 && x <= 16) forced = (uint32_t)x;
		}
		if (forced) batchedWorkers = (uint32_t)cus_ * 4 * forced;
		else
		{
			std::vector<uint32_t> sample;
			const size_t stride = std::max<size_t>(1, n / 4096);
			for (size_t k = 0; k < n; k += stride) sample.push_back(which[k]);
			std::vector<gad::ReadJob> rj;
			const uint32_t maxLen = makeJobs(sample, rj);
			std::vector<gad::ReadResult> res;
			float best = 0;
			for (uint32_t w : {2u, 3u, 4u, 5u, 6u, 8u})
			{
				const uint32_t workers = (uint32_t)cus_ * 4 * w;
				const float ms = runPersistentBatch(dg, p, rj, workers, caps, maxLen, res);
				if (batchedWorkers == 0 || ms < best)
				{
					best = ms;
					batchedWorkers = workers;
				}
			}
		}
		batchedTuneMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		out.tuneMs = batchedTuneMs;
	}
	out.workers = batchedWorkers;
	std::vector<gad::ReadJob> rj;
	const uint32_t maxLen = makeJobs(which, rj);
	std::vector<gad::ReadResult> res;
	out.kernelMs += runPersistentBatch(dg, p, rj, batchedWorkers, caps, maxLen, res);
	out.launches++;
	collectReadLoops(rj, res, which, out, bAlns_.as<uint64_t>(), bPool_.as<uint64_t>());
	// overflowed reads: the launch-per-batch form with growing workspaces
	std::vector<uint32_t> again;
	for (uint32_t id : which)
		if (out.results[id].status == gad::READLOOP_OVERFLOW) again.push_back(id);
	if (!again.empty())
	{
		std::vector<ReadLoopJob> sub;
		for (uint32_t id : again) sub.push_back(jobs[id]);
		ReadLoopOutput ro;
		runReadLoops(g, p, seeds, sub, ro);
		out.reruns += again.size() + ro.reruns;
		out.kernelMs += ro.kernelMs;
		out.launches += ro.launches;
		out.wordSteps += ro.wordSteps;
		out.extensions += ro.extensions;
		const uint64_t alnBase = out.alns.size(), poolBase = out.pool.size();
		out.alns.insert(out.alns.end(), ro.alns.begin(), ro.alns.end());
		out.pool.insert(out.pool.end(), ro.pool.begin(), ro.pool.end());
		for (size_t k = 0; k < again.size(); k++)
		{
			const uint32_t id = again[k];
			out.results[id] = ro.results[k];
			out.alnOffset[id] = alnBase + ro.alnOffset[k];
			out.poolOffset[id] = poolBase + ro.poolOffset[k];
		}
	}
}

} // namespace ga
