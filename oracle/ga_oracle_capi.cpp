// C entry points of the CPU oracle. Test infrastructure only: loaded by tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include "ga_oracle.hpp"

using namespace orc;

static uint64_t& lastWordSteps()
{
	static uint64_t n = 0;
	return n;
}

extern "C" {

struct OrcParams
{
	int64_t bandwidth;
	int64_t maxCellsPerSlice; // <= 0 means unlimited
	int32_t errorCost;
	int32_t xdrop;
	double seedExtendDensity;
	double multimapScoreFraction;
	int32_t sloppy;
	int32_t minClusterSize;
	double minAlignmentScore;
	int32_t cigarMatchMismatchMerge;
};

static Params toParams(const OrcParams* p)
{
	Params r;
	r.alignmentBandwidth = p->bandwidth;
	r.maxCellsPerSlice = p->maxCellsPerSlice <= 0 ? NMAX : (size_t)p->maxCellsPerSlice;
	r.XscoreErrorCost = p->errorCost;
	r.Xdropcutoff = p->xdrop;
	r.seedExtendDensity = p->seedExtendDensity;
	r.multimapScoreFraction = p->multimapScoreFraction;
	r.sloppyOptimizations = p->sloppy != 0;
	r.minSeedClusterSize = p->minClusterSize;
	r.minAlignmentScore = p->minAlignmentScore;
	r.cigarMatchMismatchMerge = p->cigarMatchMismatchMerge != 0;
	return r;
}

static void setErr(char* err, int cap, const char* msg)
{
	if (err && cap > 0)
	{
		std::strncpy(err, msg, cap - 1);
		err[cap - 1] = 0;
	}
}

void* orc_graph_load(const char* path, char* err, int errcap)
{
	try
	{
		return new AlignmentGraph(LoadGraphFromGFA(path));
	}
	catch (const std::exception& e)
	{
		setErr(err, errcap, e.what());
	}
	catch (const AssertFail& a)
	{
		setErr(err, errcap, a.what);
	}
	return nullptr;
}

void orc_graph_free(void* g) { delete (AlignmentGraph*)g; }

int64_t orc_graph_num_nodes(void* g) { return ((AlignmentGraph*)g)->NodeSize(); }

int64_t orc_graph_num_edges(void* g)
{
	auto* graph = (AlignmentGraph*)g;
	int64_t e = 0;
	for (auto& v : graph->outNeighbors) e += v.size();
	return e;
}

int64_t orc_graph_first_ambiguous(void* g) { return ((AlignmentGraph*)g)->firstAmbiguous; }

// Per-node arrays; neighbour lists in CSR form, in the reference's vector order.
int orc_graph_export(void* g, int64_t* nodeLength, int64_t* nodeIDs, int64_t* nodeOffset, int64_t* comp, uint8_t* linearizable, uint8_t* reverse, int64_t* chain, int64_t* chainPos, uint64_t* seq2, int64_t* inOff, int64_t* inAdj, int64_t* outOff, int64_t* outAdj)
{
	auto* graph = (AlignmentGraph*)g;
	size_t n = graph->NodeSize();
	int64_t ie = 0, oe = 0;
	for (size_t i = 0; i < n; i++)
	{
		nodeLength[i] = graph->nodeLength[i];
		nodeIDs[i] = graph->nodeIDs[i];
		nodeOffset[i] = graph->nodeOffset[i];
		comp[i] = graph->componentNumber[i];
		linearizable[i] = graph->linearizable[i];
		reverse[i] = graph->reverse[i];
		chain[i] = graph->chainNumber[i];
		chainPos[i] = graph->chainApproxPos[i];
		if (i < graph->firstAmbiguous)
		{
			seq2[2 * i] = graph->nodeSequences[i][0];
			seq2[2 * i + 1] = graph->nodeSequences[i][1];
		}
		else
		{
			seq2[2 * i] = 0;
			seq2[2 * i + 1] = 0;
		}
		inOff[i] = ie;
		for (auto x : graph->inNeighbors[i]) inAdj[ie++] = x;
		outOff[i] = oe;
		for (auto x : graph->outNeighbors[i]) outAdj[oe++] = x;
	}
	inOff[n] = ie;
	outOff[n] = oe;
	return 0;
}

void* orc_index_build(void* g, int k, int w, double discard)
{
	return new MinimizerIndex(*(AlignmentGraph*)g, k, w, 1.0 - discard);
}

void orc_index_free(void* idx) { delete (MinimizerIndex*)idx; }

int64_t orc_index_maxcount(void* idx) { return ((MinimizerIndex*)idx)->maxCount; }

int64_t orc_index_num_kmers(void* idx) { return ((MinimizerIndex*)idx)->positions.size(); }

// Seeds of one read: 10 int64 per seed in SeedHit field order. Pass order != 0 to
// run OrderSeeds (src/GraphAligner.h:250-312) on them as well.
int orc_get_seeds(void* g, void* idx, const char* seq, int64_t len, double density, int order, int64_t* out, int64_t cap, int64_t* n)
{
	auto* index = (MinimizerIndex*)idx;
	std::string s(seq, len);
	auto seeds = index->getSeeds(s, density);
	if (order)
	{
		Params p;
		Aligner a(p, *(AlignmentGraph*)g);
		try
		{
			a.orderSeedsByChaining(seeds);
		}
		catch (const AssertFail&)
		{
			*n = 0;
			return 2;
		}
	}
	*n = seeds.size();
	if ((int64_t)seeds.size() > cap) return 1;
	for (size_t i = 0; i < seeds.size(); i++)
	{
		int64_t* o = out + 10 * i;
		o[0] = seeds[i].nodeID;
		o[1] = seeds[i].nodeOffset;
		o[2] = seeds[i].seqPos;
		o[3] = seeds[i].matchLen;
		o[4] = seeds[i].reverse;
		o[5] = seeds[i].alignmentGraphNodeId;
		o[6] = seeds[i].alignmentGraphNodeOffset;
		o[7] = seeds[i].rawSeedGoodness;
		o[8] = seeds[i].seedGoodness;
		o[9] = seeds[i].seedClusterSize;
	}
	return 0;
}

// One extension in one direction (getReverseTraceFromSeed). Returns 0 on success,
// 1 if the trace failed (no slice kept), 2 on a reference assertion, 3 if the
// output buffers are too small. Trace rows: node, nodeOffset, seqPos, nodeSwitch.
int orc_extend_scheduled(void* g, const OrcParams* p, const char* seq, int64_t len, int bigraphNodeId, int64_t offset, const int32_t* schedule, int64_t scheduleLen, int64_t* traceOut, int64_t cap, int64_t* ntrace, int32_t* score, SliceSummary* summaries, int64_t scap, int64_t* nslices);

int orc_extend(void* g, const OrcParams* p, const char* seq, int64_t len, int bigraphNodeId, int64_t offset, int64_t* traceOut, int64_t cap, int64_t* ntrace, int32_t* score, SliceSummary* summaries, int64_t scap, int64_t* nslices)
{
	return orc_extend_scheduled(g, p, seq, len, bigraphNodeId, offset, nullptr, 0, traceOut, cap, ntrace, score, summaries, scap, nslices);
}

// orc_extend with a per-slice bandwidth schedule: slice s uses
// schedule[min(s, scheduleLen - 1)] (scheduleLen 0: the fixed bandwidth).
int orc_extend_scheduled(void* g, const OrcParams* p, const char* seq, int64_t len, int bigraphNodeId, int64_t offset, const int32_t* schedule, int64_t scheduleLen, int64_t* traceOut, int64_t cap, int64_t* ntrace, int32_t* score, SliceSummary* summaries, int64_t scap, int64_t* nslices)
{
	Params params = toParams(p);
	if (schedule && scheduleLen > 0) params.bandwidthSchedule.assign(schedule, schedule + scheduleLen);
	Extender ext(params, *(AlignmentGraph*)g);
	std::string_view view(seq, len);
	std::vector<SliceSummary> sums;
	*ntrace = 0;
	*nslices = 0;
	try
	{
		auto trace = ext.getReverseTraceFromSeed(view, bigraphNodeId, offset, &sums);
		*nslices = sums.size();
		if (summaries)
		{
			if ((int64_t)sums.size() > scap) return 3;
			for (size_t i = 0; i < sums.size(); i++) summaries[i] = sums[i];
		}
		if (trace.failed())
		{
			*score = SMAX;
			return 1;
		}
		*score = trace.score;
		*ntrace = trace.trace.size();
		if ((int64_t)trace.trace.size() > cap) return 3;
		for (size_t i = 0; i < trace.trace.size(); i++)
		{
			traceOut[4 * i] = trace.trace[i].DPposition.node;
			traceOut[4 * i + 1] = trace.trace[i].DPposition.nodeOffset;
			traceOut[4 * i + 2] = (int64_t)trace.trace[i].DPposition.seqPos;
			traceOut[4 * i + 3] = trace.trace[i].nodeSwitch;
		}
		return 0;
	}
	catch (const AssertFail& a)
	{
		*nslices = sums.size();
		return 2;
	}
}

// Whole reads, one read per std::thread worker pulled from a shared counter
// (src/Aligner.cpp:777-780). GAF text of all reads is concatenated in read order
// into a malloc'd buffer (free with orc_free). status[i]: 0 aligned, 1 no seeds,
// 2 dropped by an assertion, 3 no alignment survived. Returns wall seconds.
double orc_align_reads(void* g, void* idx, const OrcParams* p, double seedDensity, int64_t nreads, const char* const* names, const char* const* seqs, const int64_t* lens, int nthreads, char** gafOut, int32_t* statusOut)
{
	Params params = toParams(p);
	auto* graph = (AlignmentGraph*)g;
	auto* index = (MinimizerIndex*)idx;
	std::vector<ReadOutcome> outcomes(nreads);
	std::atomic<int64_t> next{0};
	auto start = std::chrono::steady_clock::now();
	auto work = [&]() {
		while (true)
		{
			int64_t i = next.fetch_add(1);
			if (i >= nreads) break;
			outcomes[i] = alignRead(*graph, *index, params, seedDensity, names[i], std::string(seqs[i], lens[i]));
		}
	};
	if (nthreads <= 1)
	{
		work();
	}
	else
	{
		std::vector<std::thread> threads;
		for (int t = 0; t < nthreads; t++) threads.emplace_back(work);
		for (auto& t : threads) t.join();
	}
	auto end = std::chrono::steady_clock::now();
	size_t total = 1;
	for (auto& o : outcomes) total += o.gaf.size();
	char* buf = (char*)std::malloc(total);
	size_t pos = 0;
	for (int64_t i = 0; i < nreads; i++)
	{
		std::memcpy(buf + pos, outcomes[i].gaf.data(), outcomes[i].gaf.size());
		pos += outcomes[i].gaf.size();
		statusOut[i] = outcomes[i].status;
	}
	buf[pos] = 0;
	*gafOut = buf;
	uint64_t steps = 0;
	for (auto& o : outcomes) steps += o.wordSteps;
	lastWordSteps() = steps;
	return std::chrono::duration<double>(end - start).count();
}

// Forward column steps of the last orc_align_reads call (all reads).
int64_t orc_last_word_steps() { return (int64_t)lastWordSteps(); }

// GraphAligner::AlignOneWay (src/GraphAligner.h:94-181) with the caller's seeds
// (10 int64 per seed in SeedHit field order, -1 for an unset
// alignmentGraphNodeId/Offset), then, if select != 0, SelectAlignments and the
// output sorts of src/Aligner.cpp:504-549, and if mapq != 0
// AddMappingQualities first (src/Aligner.cpp:476). The result is text in a
// malloc'd buffer: "S status seedsExtended" (status 0, or 2 when a reference
// assertion fired), then per alignment "A start end score xscore seedGoodness
// mapq n" and a line of n trace items "node,offset,seqPos,nodeSwitch,seqChar,graphChar".
int orc_align_one_way(void* g, const OrcParams* p, const char* seq, int64_t len, const int64_t* seeds, int64_t nseeds, int select, int mapq, char** out)
{
	Params params = toParams(p);
	Aligner aligner(params, *(AlignmentGraph*)g);
	std::vector<SeedHit> hits(nseeds);
	for (int64_t i = 0; i < nseeds; i++)
	{
		const int64_t* x = seeds + 10 * i;
		hits[i].nodeID = (int)x[0];
		hits[i].nodeOffset = x[1];
		hits[i].seqPos = x[2];
		hits[i].matchLen = x[3];
		hits[i].reverse = x[4] != 0;
		hits[i].alignmentGraphNodeId = x[5] < 0 ? NMAX : (size_t)x[5];
		hits[i].alignmentGraphNodeOffset = x[6] < 0 ? NMAX : (size_t)x[6];
		hits[i].rawSeedGoodness = x[7];
		hits[i].seedGoodness = x[8];
		hits[i].seedClusterSize = x[9];
	}
	std::string text;
	std::string sequence(seq, len);
	int status = 0;
	AlignmentResult result;
	try
	{
		result = aligner.AlignOneWay(sequence, hits);
		if (mapq) AddMappingQualities(result.alignments);
	}
	catch (const AssertFail&)
	{
		status = 2;
	}
	if (status == 0 && select)
	{
		if (result.alignments.size() > 0) result.alignments = SelectAlignments(result.alignments, params.minAlignmentScore, params.multimapScoreFraction);
		std::sort(result.alignments.begin(), result.alignments.end(), [](const AlignmentItem& l, const AlignmentItem& r) { return l.alignmentStart < r.alignmentStart; });
		std::sort(result.alignments.begin(), result.alignments.end(), [](const AlignmentItem& l, const AlignmentItem& r) { return l.alignmentXScore > r.alignmentXScore; });
	}
	char buf[256];
	std::snprintf(buf, sizeof(buf), "S %d %zu\n", status, status == 0 ? result.seedsExtended : (size_t)0);
	text += buf;
	if (status == 0)
		for (const auto& a : result.alignments)
		{
			const auto& t = a.trace->trace;
			std::snprintf(buf, sizeof(buf), "A %zu %zu %zu %.17g %zu %d %zu\n", a.alignmentStart, a.alignmentEnd, a.alignmentScore, a.alignmentXScore, a.seedGoodness, a.mappingQuality, t.size());
			text += buf;
			for (size_t k = 0; k < t.size(); k++)
			{
				std::snprintf(buf, sizeof(buf), "%s%zu,%zu,%zu,%d,%d,%d", k ? " " : "", t[k].DPposition.node, t[k].DPposition.nodeOffset, t[k].DPposition.seqPos, (int)t[k].nodeSwitch, (int)(unsigned char)t[k].sequenceCharacter, (int)(unsigned char)t[k].graphCharacter);
				text += buf;
			}
			text += "\n";
		}
	*out = (char*)std::malloc(text.size() + 1);
	std::memcpy(*out, text.c_str(), text.size() + 1);
	return status;
}

// OrderSeeds (src/GraphAligner.h:250-312) on n seeds in the layout above, in
// place; 2 when a reference assertion fires.
int orc_order_seeds(void* g, int64_t* seeds, int64_t n)
{
	std::vector<SeedHit> hits(n);
	for (int64_t i = 0; i < n; i++)
	{
		const int64_t* x = seeds + 10 * i;
		hits[i].nodeID = (int)x[0];
		hits[i].nodeOffset = x[1];
		hits[i].seqPos = x[2];
		hits[i].matchLen = x[3];
		hits[i].reverse = x[4] != 0;
		hits[i].alignmentGraphNodeId = x[5] < 0 ? NMAX : (size_t)x[5];
		hits[i].alignmentGraphNodeOffset = x[6] < 0 ? NMAX : (size_t)x[6];
		hits[i].rawSeedGoodness = x[7];
		hits[i].seedGoodness = x[8];
		hits[i].seedClusterSize = x[9];
	}
	Params p;
	Aligner a(p, *(AlignmentGraph*)g);
	try
	{
		a.orderSeedsByChaining(hits);
	}
	catch (const AssertFail&)
	{
		return 2;
	}
	for (int64_t i = 0; i < n; i++)
	{
		int64_t* o = seeds + 10 * i;
		o[0] = hits[i].nodeID;
		o[1] = hits[i].nodeOffset;
		o[2] = hits[i].seqPos;
		o[3] = hits[i].matchLen;
		o[4] = hits[i].reverse;
		o[5] = hits[i].alignmentGraphNodeId == NMAX ? -1 : (int64_t)hits[i].alignmentGraphNodeId;
		o[6] = hits[i].alignmentGraphNodeOffset == NMAX ? -1 : (int64_t)hits[i].alignmentGraphNodeOffset;
		o[7] = hits[i].rawSeedGoodness;
		o[8] = hits[i].seedGoodness;
		o[9] = hits[i].seedClusterSize;
	}
	return 0;
}

// iterateMinimizersReal (src/MinimizerSeeder.cpp:104-189) over one string:
// (position, k-mer) pairs in callback order, at most cap of them; returns the count.
int64_t orc_minimizers(const char* str, int64_t len, int k, int w, int64_t* out, int64_t cap)
{
	std::string s(str, len);
	int64_t n = 0;
	iterateMinimizersReal(s, k, w, [&](size_t pos, size_t kmer) {
		if (n < cap)
		{
			out[2 * n] = (int64_t)pos;
			out[2 * n + 1] = (int64_t)kmer;
		}
		n++;
	});
	return n;
}

// The index's table: distinct k-mers (in the table's iteration order), the number
// of positions of each, and the positions (splitNode << 6 | offset) of each in
// storage order, concatenated. Returns the number of positions in all.
int64_t orc_index_table(void* idx, uint64_t* kmers, int64_t* counts, uint64_t* positions)
{
	auto* index = (MinimizerIndex*)idx;
	int64_t at = 0, k = 0;
	for (const auto& p : index->positions)
	{
		if (kmers) kmers[k] = p.first;
		if (counts) counts[k] = (int64_t)p.second.size();
		for (uint64_t v : p.second)
		{
			if (positions) positions[at] = v;
			at++;
		}
		k++;
	}
	return at;
}

// WordSlice primitives, exposed for property tests against brute force.
void orc_merge(uint64_t lvp, uint64_t lvn, int32_t lend, uint64_t rvp, uint64_t rvn, int32_t rend, uint64_t* vp, uint64_t* vn, int32_t* end)
{
	auto r = WordSlice{lvp, lvn, lend}.mergeWith(WordSlice{rvp, rvn, rend});
	*vp = r.VP;
	*vn = r.VN;
	*end = r.scoreEnd;
}

void orc_difference_masks(uint64_t lvp, uint64_t lvn, uint64_t rvp, uint64_t rvn, int32_t diff, uint64_t* leftSmaller, uint64_t* rightSmaller)
{
	auto r = WordSlice::differenceMasksBitTwiddle(lvp, lvn, rvp, rvn, diff);
	*leftSmaller = r.first;
	*rightSmaller = r.second;
}

int32_t orc_max_xscore(uint64_t vp, uint64_t vn, int32_t end, int32_t seqOffset, int32_t cost, int32_t cells, uint64_t extraMask)
{
	return WordSlice{vp, vn, end}.maxXScoreLocalMinima(seqOffset, cost, cells, extraMask);
}

int32_t orc_changed_min_score(uint64_t vp, uint64_t vn, int32_t end, uint64_t ovp, uint64_t ovn, int32_t oend)
{
	return WordSlice{vp, vn, end}.changedMinScore(WordSlice{ovp, ovn, oend});
}

void orc_next_slice(uint64_t eq, uint64_t vp, uint64_t vn, int32_t end, uint64_t hinP, uint64_t hinN, uint64_t* ovp, uint64_t* ovn, int32_t* oend, uint64_t* ph, uint64_t* mh)
{
	Word hp = hinP, hn = hinN, Ph, Mh;
	auto r = getNextSlice(eq, WordSlice{vp, vn, end}, hp, hn, Ph, Mh);
	*ovp = r.VP;
	*ovn = r.VN;
	*oend = r.scoreEnd;
	*ph = Ph;
	*mh = Mh;
}

void orc_free(void* p) { std::free(p); }

// Counters of the reference's debug-only invariant checks (builds with
// ORC_EXTRA_CHECKS; -1 otherwise): checks run and checks violated.
int64_t orc_extra_checks_run()
{
#ifdef ORC_EXTRA_CHECKS
	return (int64_t)extraChecksRun();
#else
	return -1;
#endif
}

int64_t orc_extra_check_failures()
{
#ifdef ORC_EXTRA_CHECKS
	return (int64_t)extraCheckFailures();
#else
	return -1;
#endif
}

int64_t orc_slice_summary_size() { return sizeof(SliceSummary); }
}
