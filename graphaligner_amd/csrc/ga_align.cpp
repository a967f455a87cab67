// Whole-read alignment on one GPU (see ga_align.hpp).
#include "ga_align.hpp"
#include "ga_proposal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <type_traits>

namespace ga {

namespace {

// A reference assertion on the read's path; the read is dropped
// (src/Aligner.cpp:493-500).
struct ReadAssert
{
};

inline void check(bool c)
{
	if (!c) throw ReadAssert{};
}

constexpr uint64_t kNoPos = ~0ull;

// Bases an IUPAC letter stands for (A 1, C 2, G 4, T/U 8); '-' stands for none,
// letters outside the table for -1 (src/CommonUtils.cpp:67-134).
int baseSet(char c)
{
	switch (c)
	{
		case '-': return 0;
		case 'A': case 'a': return 1;
		case 'C': case 'c': return 2;
		case 'G': case 'g': return 4;
		case 'T': case 't': case 'U': case 'u': return 8;
		case 'N': case 'n': return 15;
		case 'R': case 'r': return 1 | 4;
		case 'Y': case 'y': return 2 | 8;
		case 'K': case 'k': return 4 | 8;
		case 'M': case 'm': return 2 | 1;
		case 'S': case 's': return 2 | 4;
		case 'W': case 'w': return 1 | 8;
		case 'B': case 'b': return 2 | 4 | 8;
		case 'D': case 'd': return 1 | 4 | 8;
		case 'H': case 'h': return 1 | 2 | 8;
		case 'V': case 'v': return 1 | 2 | 4;
	}
	return -1;
}

// baseSet of every byte, for the per-item checks of the trace rebuild and GAF
struct BaseSetTable
{
	int8_t v[256];
	BaseSetTable()
	{
		for (int c = 0; c < 256; c++) v[c] = (int8_t)baseSet((char)c);
	}
};
const BaseSetTable kBaseSets;

// src/GraphAlignerCommon.h:176-279 (characterMatch / ambiguousMatch): an exact
// read base matches any graph letter standing for it; otherwise the first base
// (in A, C, G, T order) both letters stand for decides, so the letters match when
// their base sets meet. A letter outside the tables fails the reference's check.
inline bool characterMatch(char s, char g)
{
	if (s == g) return true;
	if (s == '-') return false;
	const int ss = kBaseSets.v[(unsigned char)s];
	check(ss >= 0);
	const int gs = kBaseSets.v[(unsigned char)g];
	check(gs >= 0);
	return (ss & gs) != 0;
}

// CommonUtils::Complement (src/CommonUtils.cpp:67-97); 0 for letters outside the table.
char complementOf(char c)
{
	switch (c)
	{
		case 'A': case 'a': return 'T';
		case 'C': case 'c': return 'G';
		case 'T': case 't': return 'A';
		case 'G': case 'g': return 'C';
		case 'N': case 'n': return 'N';
		case 'U': case 'u': return 'A';
		case 'R': case 'r': return 'Y';
		case 'Y': case 'y': return 'R';
		case 'K': case 'k': return 'M';
		case 'M': case 'm': return 'K';
		case 'S': case 's': return 'S';
		case 'W': case 'w': return 'W';
		case 'B': case 'b': return 'V';
		case 'V': case 'v': return 'B';
		case 'D': case 'd': return 'H';
		case 'H': case 'h': return 'D';
	}
	return 0;
}

// whether complementOf knows each byte value, for the per-read letter check
struct ComplementTable
{
	uint8_t known[256];
	ComplementTable()
	{
		for (int c = 0; c < 256; c++) known[c] = complementOf((char)c) != 0 ? 1 : 0;
	}
};
const ComplementTable kComplementable;

// every letter of s[0, n) known to complementOf (a branch-free pass over the read)
bool allComplementable(const char* s, uint64_t n)
{
	uint8_t all = 1;
	for (uint64_t k = 0; k < n; k++) all &= kComplementable.known[(unsigned char)s[k]];
	return all != 0;
}

// Runs fn(0) .. fn(n - 1) on up to `threads` threads. An exception thrown by fn
// stops the hand-out of further indices; every worker is joined and the first
// exception is rethrown on the calling thread (so callers' handlers see it as in
// a single-threaded loop and the process is never terminated by a worker).
template <typename F>
void parallelFor(size_t n, int threads, F fn)
{
	if (n == 0) return;
	int t = std::max(1, std::min<int>(threads, (int)std::min<size_t>(n, 64)));
	if (t == 1)
	{
		for (size_t i = 0; i < n; i++) fn(i);
		return;
	}
	std::atomic<size_t> next{0};
	std::atomic<bool> failed{false};
	std::mutex errorMutex;
	std::exception_ptr error;
	auto work = [&]() {
		while (!failed.load(std::memory_order_relaxed))
		{
			size_t i = next.fetch_add(1);
			if (i >= n) break;
			try
			{
				fn(i);
			}
			catch (...)
			{
				std::lock_guard<std::mutex> lock(errorMutex);
				if (!error) error = std::current_exception();
				failed.store(true);
				break;
			}
		}
	};
	std::vector<std::thread> pool;
	pool.reserve(t);
	try
	{
		for (int k = 0; k < t; k++) pool.emplace_back(work);
	}
	catch (...)
	{
		// a thread that could not be started: the ones running finish the work
		if (pool.empty()) throw;
	}
	for (auto& th : pool) th.join();
	if (error) std::rethrow_exception(error);
}

double msSince(std::chrono::steady_clock::time_point t0)
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Clears an outcome for a new alignment, keeping its text buffer: callers that
// align batch after batch reuse the buffers instead of releasing hundreds of MB
// of text per batch.
void resetOutcome(ReadOutcome& o)
{
	o.status = READ_NO_ALIGNMENT;
	o.gaf.clear();
	o.seedsExtended = 0;
}

// One device trace item (packTrace): split node, offset in it, read position
// (kNoPos for the seed column, stored as -1) and the node-switch flag.
struct RawStep
{
	uint32_t node;
	uint32_t off;
	uint64_t seqPos;
	bool sw;
};

RawStep unpackRaw(uint64_t v)
{
	const uint32_t hi = (uint32_t)(v >> 32);
	RawStep r;
	r.node = (uint32_t)v;
	r.off = (hi >> 1) & 63;
	r.seqPos = (uint64_t)((int64_t)(hi >> 7) - 1);
	r.sw = (hi & 1) != 0;
	return r;
}

// GraphAligner::verifyTrace (src/GraphAligner.h:746-766) for one step o -> n of a
// split-node trace in alignment order.
void verifyStep(const GraphHost& g, const RawStep& o, const RawStep& n)
{
	check(n.seqPos != o.seqPos || n.node != o.node || n.off != o.off);
	const bool seqStep = n.seqPos == o.seqPos || n.seqPos == o.seqPos + 1;
	if (o.node == n.node)
	{
		const bool inside = (n.off == o.off || n.off == o.off + 1) && seqStep;
		const bool wrap = o.off == g.nodeLen[n.node] - 1 && n.off == 0 && seqStep;
		check(inside || wrap);
		return;
	}
	bool edge = false;
	for (uint32_t e = g.outOff[o.node]; e < g.outOff[o.node + 1]; e++)
		if (g.outAdj[e] == n.node) edge = true;
	check(edge);
	check(seqStep);
	check(o.off == g.nodeLen[o.node] - 1);
	check(n.off == 0);
}

// One term of the fixed-up trace's edit score (src/GraphAligner.h:540-557): item
// t after item p, or the first item of its direction when p is null.
int32_t editTerm(const TItem* p, const TItem& t)
{
	if (!p) return characterMatch(t.seqChar, t.graphChar) ? 0 : 1;
	if (t.seqPos == p->seqPos) return 1;
	if (!p->nodeSwitch && t.node == p->node && t.nodeOffset == p->nodeOffset) return 1;
	return characterMatch(t.seqChar, t.graphChar) ? 0 : 1;
}

// AlignmentSelection::alignmentIncompatible (src/AlignmentSelection.cpp:13-30)
bool alignmentIncompatible(const Alignment& left, const Alignment& right)
{
	const float cutoff = 0.05;
	auto minOverlapLen = std::min(left.alignmentEnd - left.alignmentStart, right.alignmentEnd - right.alignmentStart) * cutoff;
	uint64_t ls = left.alignmentStart, le = left.alignmentEnd, rs = right.alignmentStart, re = right.alignmentEnd;
	if (ls > rs)
	{
		std::swap(ls, rs);
		std::swap(le, re);
	}
	int overlap = 0;
	if (le > rs) overlap = (int)(le - rs);
	return overlap > minOverlapLen;
}

} // namespace

// GraphAligner::exactAlignmentPart (src/GraphAligner.h:403-457): whether the
// alignment's trace passes through the seed's node and offset at its read
// position. A trace's read positions never decrease and rise by at most one per
// item (the step rule buildAlignment checks on every item), so each position from
// the first to the last is held by one contiguous group of items. The
// reference's bisection finds an item of the seed position's group and scans the
// group, and its range checks cannot fail on such a trace; the answer is whether
// an item of the group has the seed's node and offset. The group's first item
// lies in [p - first, p - first + z], z the trace's items that do not advance the
// read, so it is bisected there.
bool exactAlignmentPart(const Alignment& aln, const SeedHit& seed)
{
	const std::vector<TItem>& trace = aln.trace;
	check(trace.back().seqPos > trace[0].seqPos);
	if (trace.back().seqPos < seed.seqPos) return false;
	if (trace[0].seqPos > seed.seqPos) return false;
	const size_t n = trace.size();
	const uint64_t target = seed.seqPos, front = trace[0].seqPos, back = trace.back().seqPos;
	const uint64_t still = (uint64_t)(n - 1) - (back - front);
	size_t lo = (size_t)(target - front);
	size_t hi = (size_t)std::min<uint64_t>(n - 1, target - front + still);
	while (lo < hi)
	{
		const size_t mid = lo + (hi - lo) / 2;
		if (trace[mid].seqPos < target) lo = mid + 1;
		else hi = mid;
	}
	const uint64_t compareNode = (uint64_t)(int64_t)(seed.nodeID * 2 + (seed.reverse ? 1 : 0));
	for (size_t i = lo; i < n && trace[i].seqPos == target; i++)
		if (compareNode == trace[i].node && seed.nodeOffset == trace[i].nodeOffset) return true;
	return false;
}


// AlignmentSelection::SelectAlignments (src/AlignmentSelection.cpp:31-53, 65-95).
// The score threshold passes through SelectionOptions' int field
// (src/AlignmentSelection.h:21, src/Aligner.cpp:402) before the comparison.
std::vector<Alignment> selectAlignments(std::vector<Alignment>& all, double minScore, double fraction)
{
	std::vector<Alignment> kept;
	const int32_t threshold = selectionScoreThreshold(minScore);
	if (threshold > INT32_MIN)
	{
		for (auto& a : all)
			if (a.alignmentXScore >= threshold) kept.push_back(std::move(a));
	}
	else
		for (auto& a : all) kept.push_back(std::move(a));
	if (fraction == 0) return kept;
	std::vector<bool> skip(kept.size(), false);
	for (size_t i = 0; i < kept.size(); i++)
	{
		for (size_t j = 0; j < kept.size(); j++)
		{
			if (i == j) continue;
			if (!alignmentIncompatible(kept[i], kept[j])) continue;
			if (kept[i].alignmentXScore < kept[j].alignmentXScore * fraction)
			{
				skip[i] = true;
				break;
			}
		}
	}
	std::vector<Alignment> result;
	for (size_t i = 0; i < kept.size(); i++)
		if (!skip[i]) result.push_back(std::move(kept[i]));
	return result;
}

void sortForOutput(std::vector<Alignment>& alns)
{
	std::sort(alns.begin(), alns.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentStart < r.alignmentStart; });
	std::sort(alns.begin(), alns.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentXScore > r.alignmentXScore; });
}

bool addMappingQualities(std::vector<Alignment>& alns)
{
	try
	{
		for (size_t i = 0; i < alns.size(); i++)
		{
			check(alns[i].alignmentXScore != -1);
			// sum over incompatible alignments j of 10^(X_j - X_i), capped at 10
			double otherSum = 0;
			for (size_t j = 0; j < alns.size(); j++)
			{
				if (i == j || !alignmentIncompatible(alns[i], alns[j])) continue;
				check(alns[j].alignmentXScore != -1);
				if (alns[j].alignmentXScore >= alns[i].alignmentXScore + 1)
				{
					otherSum += 10;
					break;
				}
				check(alns[j].alignmentXScore <= alns[i].alignmentXScore + 1);
				otherSum += std::pow(10.0, alns[j].alignmentXScore - alns[i].alignmentXScore);
				if (otherSum >= 10) break;
			}
			int q;
			if (otherSum >= 10) q = 0;
			else if (otherSum <= 0.000001) q = 60;
			else
			{
				// int conversion of the phred value, as the reference's int field takes it
				q = (int)(-std::log(1.0 - 1.0 / (1.0 + otherSum)) * 10);
				if (q >= 60) q = 60;
			}
			check(q >= 0 && q <= 60);
			alns[i].mappingQuality = q;
		}
		return true;
	}
	catch (const ReadAssert&)
	{
		return false;
	}
}

namespace {

// decimal text of v, as std::ostream prints it, appended to s
void appendUint(std::string& s, uint64_t v)
{
	char buf[20];
	int n = 0;
	do
	{
		buf[n++] = (char)('0' + v % 10);
		v /= 10;
	} while (v != 0);
	while (n > 0) s.push_back(buf[--n]);
}

void appendInt(std::string& s, int64_t v)
{
	if (v < 0)
	{
		s.push_back('-');
		appendUint(s, (uint64_t)0 - (uint64_t)v);
	}
	else appendUint(s, (uint64_t)v);
}

// decimal digits of v written at p (as std::ostream prints it); returns the end
inline char* putUint(char* p, uint64_t v)
{
	char buf[20];
	int n = 0;
	do
	{
		buf[n++] = (char)('0' + v % 10);
		v /= 10;
	} while (v != 0);
	while (n > 0) *p++ = buf[--n];
	return p;
}

// GAF line of one alignment (src/GraphAlignerGAFAlignment.h:37-250).
// Appends the GAF line (without its newline) to s.
void appendGafLine(std::string& s, const GraphHost& g, const char* name, uint64_t readLen, const Alignment& a, bool merge)
{
	enum { Match = 0, Mismatch = 1, MatchOrMismatch = 2, Insertion = 3, Deletion = 4, Empty = 5 };
	static const char ops[] = {'=', 'X', 'M', 'I', 'D'};
	const TItem* t = a.trace.data();
	const size_t n = a.trace.size();
	// scratch kept per thread: the line's parts are built, then appended. A cigar
	// run of L >= 1 items takes 1 + digits(L) <= 2 L characters, so the cigar of
	// n items fits in 2 n bytes.
	thread_local std::string path;
	thread_local std::vector<char> cigarBuf;
	path.clear();
	if (cigarBuf.size() < 2 * n + 32) cigarBuf.resize(2 * n + 32);
	char* cg = cigarBuf.data();
	// items per edit type; without merge every item's type is its run's, so the
	// counts are added per run, with merge matches and mismatches per item
	uint64_t perType[5] = {0, 0, 0, 0, 0};
	uint64_t pathLen = 0;
	int curId = (int)(int64_t)t[0].node;
	bool curRev = (t[0].node % 2) == 1;
	uint64_t curOff = t[0].nodeOffset;
	auto addNode = [&](int id, bool rev) {
		path.push_back(rev ? '<' : '>');
		const std::string* nm = g.bigName(id);
		if (!nm) appendInt(path, id / 2);
		else path += *nm;
	};
	const bool firstMatch = characterMatch(t[0].seqChar, t[0].graphChar);
	int currentEdit = merge ? MatchOrMismatch : (firstMatch ? Match : Mismatch);
	uint64_t editLength = 1;
	uint64_t mergedMatches = firstMatch ? 1 : 0, mergedMismatches = firstMatch ? 0 : 1;
	addNode(curId, curRev);
	pathLen += g.bigSize(curId);
	for (size_t pos = 1; pos < n; pos++)
	{
		const TItem& p = t[pos - 1];
		const TItem& c = t[pos];
		bool insideNode = !p.nodeSwitch;
		if (!insideNode)
		{
			const int newId = (int)(int64_t)c.node;
			const bool newRev = (c.node % 2) == 1;
			const uint64_t newOff = c.nodeOffset;
			insideNode = newId == curId && newRev == curRev && newOff > curOff;
			if (!insideNode)
			{
				const uint64_t skippedBefore = g.bigSize(curId) - 1 - p.nodeOffset;
				curId = newId;
				curRev = newRev;
				curOff = newOff;
				addNode(curId, curRev);
				const uint64_t skippedAfter = c.nodeOffset;
				pathLen += g.bigSize(curId) - (skippedBefore + skippedAfter);
			}
		}
		int nextEdit;
		if (p.seqPos == c.seqPos) nextEdit = Deletion;
		else if (insideNode && p.nodeOffset == c.nodeOffset) nextEdit = Insertion;
		else
		{
			const bool match = characterMatch(c.seqChar, c.graphChar);
			if (merge)
			{
				nextEdit = MatchOrMismatch;
				mergedMatches += match ? 1 : 0;
				mergedMismatches += match ? 0 : 1;
			}
			else nextEdit = match ? Match : Mismatch;
		}
		if (currentEdit != nextEdit)
		{
			perType[currentEdit] += editLength;
			cg = putUint(cg, editLength);
			*cg++ = ops[currentEdit];
			currentEdit = nextEdit;
			editLength = 0;
		}
		editLength++;
	}
	perType[currentEdit] += editLength;
	cg = putUint(cg, editLength);
	*cg++ = ops[currentEdit];
	const uint64_t matches = merge ? mergedMatches : perType[Match];
	const uint64_t mismatches = merge ? mergedMismatches : perType[Mismatch];
	const uint64_t deletions = perType[Deletion], insertions = perType[Insertion];
	const size_t cigarLen = (size_t)(cg - cigarBuf.data());
	uint64_t pathEnd = pathLen - (g.bigSize((int)(int64_t)t[n - 1].node) - 1 - t[n - 1].nodeOffset);
	double total = (double)(matches + mismatches + deletions + insertions);
	s.reserve(s.size() + cigarLen + path.size() + std::strlen(name) + 192);
	s += name;
	s += '\t';
	appendUint(s, readLen);
	s += '\t';
	appendUint(s, t[0].seqPos);
	s += '\t';
	appendUint(s, t[n - 1].seqPos + 1);
	s += "\t+\t";
	s += path;
	s += '\t';
	appendUint(s, pathLen);
	s += '\t';
	appendUint(s, t[0].nodeOffset);
	s += '\t';
	appendUint(s, pathEnd);
	s += '\t';
	appendUint(s, matches);
	s += '\t';
	appendUint(s, n);
	s += '\t';
	appendInt(s, a.mappingQuality);
	s += "\tNM:i:";
	appendUint(s, mismatches + deletions + insertions);
	// floating-point fields in std::ostream's default format
	std::ostringstream f;
	if (a.alignmentXScore != -1) f << "\tAS:f:" << a.alignmentXScore;
	f << "\tdv:f:" << 1.0 - ((double)matches / total);
	f << "\tid:f:" << ((double)matches / total);
	s += f.str();
	s += "\tcg:Z:";
	s.append(cigarBuf.data(), cigarLen);
}

} // namespace

bool gafLine(std::string& s, const GraphHost& g, const char* name, uint64_t readLen, const Alignment& a, bool cigarMatchMismatchMerge)
{
	const size_t keep = s.size();
	try
	{
		appendGafLine(s, g, name, readLen, a, cigarMatchMismatchMerge);
		return true;
	}
	catch (const ReadAssert&)
	{
		s.resize(keep);
		return false;
	}
}

namespace {

// One seed of a round with its device jobs (-1: none).
struct Spec
{
	size_t seed;
	int64_t bwdJob, fwdJob;
};

// Seed loop state of one read (GraphAligner::AlignOneWay, src/GraphAligner.h:94-181).
struct ReadState
{
	uint32_t read = 0;
	uint64_t readLen = 0;
	std::vector<SeedHit> seeds;
	size_t next = 0;
	std::vector<Alignment> alignments;
	uint64_t seedScoreForEndToEndAln = 0;
	uint64_t worstExtendedSeedScore = 0;
	uint64_t seedsExtended = 0;
	uint64_t extendSeeds = 0;
	bool done = false;
	int status = READ_ALIGNED;
	// seed the loop picked last and, in the host-driven rounds, whether it is
	// still to be extended
	size_t pending = 0;
	int64_t bwdJob = -1, fwdJob = -1;
	bool pendingReady = false;
	bool finished = false; // selection and GAF done
	// seeds extended in the current round (the picked one first) and the read's
	// rounds so far
	std::vector<Spec> specs;
	// the extension jobs of specs in order (each spec's backward job, then its
	// forward job), with the specs' job indexes local to this list until the
	// round's launch list is put together
	std::vector<ExtJob> jobs;
	uint32_t roundsRun = 0;
	// outcome records: cluster of every seed and the clusters' features
	std::vector<uint32_t> clusterOf;
	std::vector<float> features;
	uint32_t nclusters = 0;
};

struct Job
{
	uint32_t state;
	bool backward;
};

} // namespace

std::vector<SeedHit> seedsFromMatches(const GraphHost& g, const MinimizerIndex& idx, const gad::SeedMatch* matches, uint32_t nmatches, uint64_t readLen, double density)
{
	std::vector<gad::SeedMatch> order(matches, matches + nmatches);
	// same comparator and input order as the reference's sort of match indices
	std::sort(order.begin(), order.end(), [](const gad::SeedMatch& l, const gad::SeedMatch& r) { return l.count < r.count; });
	uint64_t maxHits = density == -1 ? kNoPos : (uint64_t)(readLen * density);
	std::vector<SeedHit> result;
	uint64_t seedsHere = 0, allowedCount = 0;
	for (const auto& m : order)
		if ((uint64_t)m.start + m.count > idx.positions.size()) throw std::runtime_error("seed match outside the index's position array");
	{
		// the seeds the loop below takes, counted first so the result is allocated once
		uint64_t total = 0, here = 0, allowed = 0;
		for (const auto& m : order)
		{
			if (here >= maxHits && m.count > allowed) break;
			allowed = m.count;
			total += m.count;
			here += m.count;
		}
		result.reserve(total);
	}
	for (const auto& m : order)
	{
		uint64_t count = m.count;
		if (seedsHere >= maxHits && count > allowedCount) break;
		allowedCount = count;
		for (uint32_t k = 0; k < m.count; k++)
		{
			uint64_t merged = idx.positions[m.start + k];
			uint64_t split = merged >> 6;
			uint64_t offset = merged & 63;
			SeedHit h;
			h.nodeID = g.nodeId[split] / 2;
			h.nodeOffset = offset + g.nodeOff[split];
			h.seqPos = m.pos;
			h.matchLen = idx.k;
			h.rawSeedGoodness = idx.maxCount - count;
			h.reverse = g.reverse[split] != 0;
			h.alignmentGraphNodeId = split;
			h.alignmentGraphNodeOffset = offset;
			result.push_back(h);
		}
		seedsHere += count;
	}
	return result;
}

namespace {

// GraphAligner::orderSeedsByChaining (src/GraphAligner.h:250-312) with the
// reference's assertions as ReadAssert.
void orderSeedsChecked(const GraphHost& g, std::vector<SeedHit>& seeds)
{
	// seeds keyed by chain and approximate diagonal, sorted by both: each chain's
	// seeds are then consecutive in diagonal order. Clusters (diagonals within 100
	// of the previous one) and their covered bases do not depend on the order of
	// equal keys, and a failing check drops the read whichever fails first.
	struct ChainSeed
	{
		uint64_t chain, diagonal;
		uint32_t seed;
	};
	std::vector<ChainSeed> keys(seeds.size());
	for (size_t i = 0; i < seeds.size(); i++)
	{
		uint64_t node = seeds[i].alignmentGraphNodeId;
		uint64_t off = seeds[i].alignmentGraphNodeOffset;
		if (node == kNoPos)
		{
			// a seed from outside the minimizer seeder: its split node from the
			// bigraph position
			const int forwardId = seeds[i].reverse ? seeds[i].nodeID * 2 + 1 : seeds[i].nodeID * 2;
			node = g.unitigNode(forwardId, seeds[i].nodeOffset);
			check(seeds[i].nodeOffset >= g.nodeOff[node]);
			off = seeds[i].nodeOffset - g.nodeOff[node];
		}
		check(g.chainPos[node] + off >= seeds[i].seqPos);
		keys[i] = ChainSeed{g.chain[node], g.chainPos[node] + off - seeds[i].seqPos, (uint32_t)i};
	}
	std::sort(keys.begin(), keys.end(), [](const ChainSeed& l, const ChainSeed& r) { return l.chain != r.chain ? l.chain < r.chain : l.diagonal < r.diagonal; });
	size_t clusterStart = 0;
	for (size_t i = 1; i <= keys.size(); i++)
	{
		if (i < keys.size() && keys[i].chain == keys[i - 1].chain && keys[i].diagonal <= keys[i - 1].diagonal + 100) continue;
		check(i > clusterStart);
		std::sort(keys.begin() + clusterStart, keys.begin() + i, [&seeds](const ChainSeed& l, const ChainSeed& r) { return seeds[l.seed].seqPos < seeds[r.seed].seqPos; });
		uint64_t matchingBps = 0;
		int lastEnd = INT_MIN;
		for (size_t k = clusterStart; k < i; k++)
		{
			const SeedHit& s = seeds[keys[k].seed];
			int thisStart = (int)s.seqPos - (int)s.matchLen + 1;
			int thisEnd = (int)s.seqPos;
			check(thisEnd >= lastEnd);
			check(thisEnd > thisStart);
			matchingBps += (thisEnd - std::max(thisStart, lastEnd));
			lastEnd = thisEnd;
		}
		for (size_t k = clusterStart; k < i; k++)
		{
			seeds[keys[k].seed].seedGoodness = matchingBps + seeds[keys[k].seed].rawSeedGoodness;
			seeds[keys[k].seed].seedClusterSize = i - clusterStart;
		}
		clusterStart = i;
	}
	// the reference's std::sort by goodness, then its reverse: std::sort's steps
	// depend only on the comparisons, so sorting (goodness, index) pairs with the
	// same comparator gives the same permutation of the seeds, ties included,
	// while moving 16-byte pairs instead of whole seeds
	std::vector<std::pair<uint64_t, uint32_t>> byGoodness(seeds.size());
	for (size_t i = 0; i < seeds.size(); i++) byGoodness[i] = {seeds[i].seedGoodness, (uint32_t)i};
	std::sort(byGoodness.begin(), byGoodness.end(), [](const std::pair<uint64_t, uint32_t>& l, const std::pair<uint64_t, uint32_t>& r) { return l.first < r.first; });
	std::vector<SeedHit> ordered(seeds.size());
	for (size_t j = 0; j < seeds.size(); j++) ordered[j] = seeds[byGoodness[seeds.size() - 1 - j].second];
	seeds.swap(ordered);
}

} // namespace

bool orderSeedsByChaining(const GraphHost& g, std::vector<SeedHit>& seeds)
{
	try
	{
		orderSeedsChecked(g, seeds);
		return true;
	}
	catch (const ReadAssert&)
	{
		return false;
	}
}

void seedReads(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, double density, int order, const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens, std::vector<std::vector<SeedHit>>& out, double* lookupMs, std::vector<int32_t>* status)
{
	ctx.uploadReads(seqs, lens);
	SeedMatches sm;
	ctx.lookupSeeds(idx, 0, (uint32_t)seqs.size(), sm);
	if (lookupMs) *lookupMs = sm.lookupMs;
	out.assign(seqs.size(), {});
	if (status) status->assign(seqs.size(), 0);
	if (order == 2)
	{
		std::vector<uint8_t> asserted;
		ctx.orderedSeeds(g, idx, 0, (uint32_t)seqs.size(), sm, density, out, asserted);
		if (status)
			for (size_t i = 0; i < seqs.size(); i++)
				if (asserted[i]) (*status)[i] = READ_ASSERT;
		return;
	}
	parallelFor(seqs.size(), 16, [&](size_t i) {
		out[i] = seedsFromMatches(g, idx, sm.data() + sm.matchOffset[i], sm.matchCount[i], lens[i], density);
		if (order && !orderSeedsByChaining(g, out[i]) && status) (*status)[i] = READ_ASSERT;
	});
}

namespace {

// getAlignmentFromSeed (src/GraphAligner.h:631-692) rebuilt from the two raw
// traces the device seed loop kept (either may be empty: that direction failed
// or did not run): the checks of getTwoDirectionalTrace, verifyTrace, the
// fix-ups with their score recount, the merge and the X-score.
// Sequential readers of one direction's trace (end first): packed items as the
// device wrote them, or the compact codes of ga_trace_codes.h.
struct PackedReader
{
	const uint64_t* p;
	uint64_t next() { return *p++; }
};

struct CodeReader
{
	const uint8_t* codes;
	const std::pair<uint64_t, uint64_t>* esc;
	const std::pair<uint64_t, uint64_t>* escEnd;
	uint64_t index;
	uint64_t prev;
	gad::TraceGraph tg;
	CodeReader(const GraphHost& g, const ExtBatch& b, size_t job)
	    : codes(b.codes), esc(nullptr), escEnd(b.escapes.data() + b.escapes.size()), index(b.traceOffset[job]), prev(0), tg{g.nodeLen.data(), g.inOff.data(), g.inAdj.data()}
	{
		esc = b.escapes.data() + (std::lower_bound(b.escapes.begin(), b.escapes.end(), std::make_pair(index, (uint64_t)0)) - b.escapes.begin());
	}
	uint64_t next()
	{
		const uint32_t c = codes[index];
		if (c == gad::kTraceEscape)
		{
			if (esc == escEnd || esc->first != index) throw std::runtime_error("trace codes: escaped item missing");
			prev = esc->second;
			++esc;
		}
		else prev = gad::traceDecode(tg, prev, c);
		index++;
		return prev;
	}
	// the code of the item next() would read
	uint32_t peek() const { return codes[index]; }
	// passes over n items the caller derived itself, the last of which is `last`
	void skip(uint32_t n, uint64_t last)
	{
		index += n;
		prev = last;
	}
};

// Code of an item one read row up and one column left of the item before it in
// the same node, without a node switch (ga_trace_codes.h): the diagonal step
// that makes up most of a trace. Such an item follows from the one before it,
// and the step passes verifyTrace's check by construction.
constexpr uint32_t kDiagonalCode = (1u << 2) | (1u << 1);

uint64_t packRaw(const RawStep& r)
{
	const uint32_t hi = ((uint32_t)(r.seqPos + 1) << 7) | (r.off << 1) | (r.sw ? 1u : 0u);
	return ((uint64_t)hi << 32) | r.node;
}

RawStep diagonalBefore(const RawStep& r)
{
	RawStep d;
	d.node = r.node;
	d.off = r.off - 1;
	d.seqPos = r.seqPos - 1;
	d.sw = false;
	return d;
}

// What the trace rebuild needs of one split node, refreshed when the trace
// enters another node: its bigraph id and offset base in either direction and
// its bases.
struct NodeView
{
	uint32_t split = UINT32_MAX;
	uint64_t bigNode = 0;
	uint64_t offBase = 0;
	uint64_t w0 = 0, w1 = 0;
	bool amb = false;
	// reverse: the bigraph node of the other strand, offsets counted from its start
	void load(const GraphHost& g, uint32_t n, bool reverse)
	{
		split = n;
		const int id = g.nodeId[n];
		if (reverse)
		{
			bigNode = (uint64_t)(int64_t)((id % 2 == 0) ? (id / 2) * 2 + 1 : (id / 2) * 2);
			offBase = g.splitBigSize[n] - g.nodeOff[n] - 1;
		}
		else
		{
			bigNode = (uint64_t)(int64_t)id;
			offBase = g.nodeOff[n];
		}
		amb = n >= g.firstAmbiguous;
		if (!amb)
		{
			w0 = g.seq2[2 * (uint64_t)n];
			w1 = g.seq2[2 * (uint64_t)n + 1];
		}
	}
	char base(const GraphHost& g, uint32_t off) const
	{
		if (amb) return g.nodeChar(split, off);
		return "ACGT"[((off < 32 ? w0 : w1) >> ((off & 31) * 2)) & 3];
	}
	// complementOf(base(g, off)); 0 for a letter without a complement
	char complementBase(const GraphHost& g, uint32_t off) const
	{
		if (amb) return complementOf(g.nodeChar(split, off));
		return "TGCA"[((off < 32 ? w0 : w1) >> ((off & 31) * 2)) & 3];
	}
};

inline void verifyFast(const GraphHost& g, const RawStep& o, const RawStep& n)
{
	if (o.node == n.node && n.off == o.off + 1 && (n.seqPos == o.seqPos || n.seqPos == o.seqPos + 1)) return;
	verifyStep(g, o, n);
}

template <typename Reader>
Alignment buildAlignment(const GraphHost& g, const AlignOptions& opt, const char* seq, uint64_t L, const SeedHit& s, Reader bwd, uint32_t nb, Reader fwd, uint32_t nf)
{
	// One pass per direction, in the order the traces are stored, builds the
	// merged trace directly: the backward trace (its order is read order after
	// fixReverseTraceSeqPosAndOrder, node-switch flags moved one item back) at
	// [0, nb), then the forward trace reversed (fixForwardTraceSeqPos) from
	// nb - 1 on, its first item replacing the backward trace's last
	// (getTwoDirectionalTrace's merge). Every step is checked as verifyTrace does
	// and the edit scores are summed term by term.
	const int forwardId = s.reverse ? s.nodeID * 2 + 1 : s.nodeID * 2;
	const int backwardId = s.reverse ? s.nodeID * 2 : s.nodeID * 2 + 1;
	const auto rp = g.reversePosition(forwardId, s.nodeOffset);
	Alignment item;
	if (nb == 0 && nf == 0) return item;
	std::vector<TItem>& out = item.trace;
	out.resize((size_t)nb + nf - (nb > 0 && nf > 0 ? 1 : 0));
	int32_t score = 0;
	NodeView view;
	constexpr bool kCodes = std::is_same<Reader, CodeReader>::value;
	if (nb > 0)
	{
		const uint64_t end = s.seqPos - 1;
		RawStep next = unpackRaw(bwd.next());
		for (uint32_t k = 0; k < nb; k++)
		{
			const RawStep cur = next;
			if (k + 1 < nb)
			{
				next = unpackRaw(bwd.next());
				verifyFast(g, next, cur);
			}
			else check(cur.seqPos == kNoPos && g.nodeId[cur.node] == backwardId && g.nodeOff[cur.node] + cur.off == rp.second);
			if (cur.node != view.split) view.load(g, cur.node, true);
			TItem& t = out[k];
			t.node = view.bigNode;
			t.nodeOffset = view.offBase - cur.off;
			t.seqPos = end - cur.seqPos;
			t.nodeSwitch = k + 1 < nb ? next.sw : false;
			check(t.seqPos < L);
			t.seqChar = seq[t.seqPos];
			const char c = view.complementBase(g, cur.off);
			check(c != 0);
			t.graphChar = c;
			score += editTerm(k == 0 ? nullptr : &out[k - 1], t);
			if constexpr (kCodes)
			{
				// the same steps for items k + 1, ... while the item after each is a
				// diagonal step (so it follows from it, passes the step check and
				// has no node switch) and the item lies in the node of the view
				uint32_t taken = 0;
				while (k + 2 < nb && next.node == view.split && bwd.peek() == kDiagonalCode)
				{
					k++;
					const RawStep item = next;
					next = diagonalBefore(item);
					taken++;
					bwd.skip(1, 0);
					TItem& u = out[k];
					u.node = view.bigNode;
					u.nodeOffset = view.offBase - item.off;
					u.seqPos = end - item.seqPos;
					u.nodeSwitch = false;
					check(u.seqPos < L);
					u.seqChar = seq[u.seqPos];
					const char d = view.complementBase(g, item.off);
					check(d != 0);
					u.graphChar = d;
					score += editTerm(&out[k - 1], u);
				}
				if (taken > 0) bwd.skip(0, packRaw(next));
			}
		}
	}
	if (nf > 0)
	{
		const uint64_t start = s.seqPos + 1;
		const size_t last = (nb > 0 ? nb - 1 : 0) + nf - 1; // where raw item 0 goes
		RawStep prevRaw{};
		TItem prevT{};
		view.split = UINT32_MAX;
		for (uint32_t k = 0; k < nf; k++)
		{
			const RawStep cur = unpackRaw(fwd.next());
			if (k > 0) verifyFast(g, cur, prevRaw);
			if (cur.node != view.split) view.load(g, cur.node, false);
			TItem t;
			t.node = view.bigNode;
			t.nodeOffset = view.offBase + cur.off;
			t.seqPos = cur.seqPos + start;
			t.nodeSwitch = cur.sw;
			check(t.seqPos < L);
			t.seqChar = seq[t.seqPos];
			t.graphChar = view.base(g, cur.off);
			// the term of the step from this item to the one after it in read order
			if (k > 0) score += editTerm(&t, prevT);
			if (k + 1 == nf)
			{
				check(cur.seqPos == kNoPos && g.nodeId[cur.node] == forwardId && g.nodeOff[cur.node] + cur.off == s.nodeOffset);
				score += editTerm(nullptr, t);
				if (nb > 0)
				{
					const TItem& x = out[nb - 1];
					check(x.node == t.node && x.nodeOffset == t.nodeOffset && x.seqPos == t.seqPos);
				}
			}
			out[last - k] = t;
			prevT = t;
			prevRaw = cur;
			if constexpr (kCodes)
			{
				// the same steps for the following items that are diagonal steps
				// (each follows from the one before, in the same node, passes the
				// step check and has no node switch), the last item excepted
				uint32_t taken = 0;
				while (k + 2 < nf && fwd.peek() == kDiagonalCode)
				{
					k++;
					const RawStep item = diagonalBefore(prevRaw);
					taken++;
					fwd.skip(1, 0);
					TItem u;
					u.node = view.bigNode;
					u.nodeOffset = view.offBase + item.off;
					u.seqPos = item.seqPos + start;
					u.nodeSwitch = false;
					check(u.seqPos < L);
					u.seqChar = seq[u.seqPos];
					u.graphChar = view.base(g, item.off);
					score += editTerm(&u, prevT);
					out[last - k] = u;
					prevT = u;
					prevRaw = item;
				}
				if (taken > 0) fwd.skip(0, packRaw(prevRaw));
			}
		}
	}
	check(!out.empty());
	uint64_t seqStart = out[0].seqPos;
	uint64_t seqEnd = out.back().seqPos;
	check(seqEnd < L);
	item.traceScore = score;
	item.alignmentScore = (uint64_t)(int64_t)item.traceScore;
	item.alignmentStart = seqStart;
	item.alignmentEnd = seqEnd + 1;
	item.alignmentXScore = (int32_t)item.alignmentLength() * 100 - opt.xscoreErrorCost * (int32_t)item.alignmentScore;
	item.alignmentXScore /= 100.0;
	return item;
}

} // namespace

void alignResidentDeviceLoop(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, const AlignOptions& opt, const char* const* names, uint32_t first, uint32_t count, ReadOutcome* out, AlignStats& stats)
{
	auto t0 = std::chrono::steady_clock::now();
	const ReadSet& rs = ctx.reads();
	if ((uint64_t)first + count > rs.size()) throw std::runtime_error("read range outside the resident reads");
	stats = AlignStats{};
	for (uint32_t i = 0; i < count; i++)
	{
		resetOutcome(out[i]);
		if (rs.length[first + i] > (1u << 24)) throw std::runtime_error("reads longer than 16 Mbp are not supported");
	}
	gad::ReadLoopParams rp;
	rp.ext = opt.ext;
	rp.minSeedClusterSize = opt.minSeedClusterSize;
	rp.sloppy = opt.sloppyOptimizations ? 1 : 0;
	rp.pad = 0;
	for (uint32_t batchStart = 0; batchStart < count; batchStart += opt.readsPerBatch)
	{
		const uint32_t bn = std::min<uint32_t>(count - batchStart, opt.readsPerBatch);
		const uint32_t base = first + batchStart;
		SeedMatches rb;
		auto tl = std::chrono::steady_clock::now();
		ctx.lookupSeeds(idx, base, bn, rb);
		stats.lookupCallMs += msSince(tl);
		stats.lookupMs += rb.lookupMs;
		auto th = std::chrono::steady_clock::now();
		// seeds, seed order and the read checks that precede the seed loop
		std::vector<std::vector<SeedHit>> seeds(bn);
		std::vector<uint8_t> ready(bn, 0);
		std::vector<uint8_t> ordered;
		if (opt.deviceSeedOrder) ctx.orderedSeeds(g, idx, base, bn, rb, opt.minimizerSeedDensity, seeds, ordered);
		parallelFor(bn, opt.hostThreads, [&](size_t i) {
			ReadOutcome& o = out[batchStart + i];
			const char* seq = rs.seq(base + i);
			const uint32_t L = rs.length[base + i];
			try
			{
				if (opt.deviceSeedOrder)
				{
					check(!ordered[i]);
					if (seeds[i].empty())
					{
						o.status = READ_NO_SEEDS;
						return;
					}
				}
				else
				{
					seeds[i] = seedsFromMatches(g, idx, rb.data() + rb.matchOffset[i], rb.matchCount[i], L, opt.minimizerSeedDensity);
					if (seeds[i].empty())
					{
						o.status = READ_NO_SEEDS;
						return;
					}
					orderSeedsChecked(g, seeds[i]);
				}
				// ReverseComplement of the read asserts on letters it does not know
				check(allComplementable(seq, L));
				ready[i] = 1;
			}
			catch (const ReadAssert&)
			{
				o.status = READ_ASSERT;
			}
		});
		std::vector<std::vector<uint32_t>> clusterOf(opt.records ? bn : 0);
		std::vector<std::vector<float>> features(opt.records ? bn : 0);
		if (opt.proposal)
		{
			std::vector<std::vector<SeedHit>*> sp;
			std::vector<const char*> sq;
			std::vector<uint32_t> ln;
			for (uint32_t i = 0; i < bn; i++)
				if (ready[i])
				{
					sp.push_back(&seeds[i]);
					sq.push_back(rs.seq(base + i));
					ln.push_back(rs.length[base + i]);
				}
			ProposalStats ps;
			applyProposal(*opt.proposal, g, idx.maxCount, sp.size(), sp.data(), sq.data(), ln.data(), opt.hostThreads, ps);
			stats.proposalClusters += ps.clusters;
			stats.prunedClusters += ps.prunedClusters;
			stats.prunedSeeds += ps.prunedSeeds;
			stats.proposalMs += ps.ms;
		}
		if (opt.records)
			parallelFor(bn, opt.hostThreads, [&](size_t i) {
				if (!ready[i]) return;
				const uint32_t nc = seedClusters(g, seeds[i], clusterOf[i]);
				features[i].resize((size_t)nc * kClusterFeatures);
				clusterFeatures(g, idx.maxCount, rs.seq(base + i), rs.length[base + i], seeds[i], clusterOf[i], nc, nc, features[i].data());
			});
		std::vector<ReadLoopJob> jobs;
		std::vector<uint32_t> jobRead;
		uint64_t totalSeeds = 0;
		for (uint32_t i = 0; i < bn; i++)
		{
			stats.seeds += seeds[i].size();
			if (!ready[i]) continue;
			ReadLoopJob j;
			j.read = base + i;
			j.seedCount = (uint32_t)seeds[i].size();
			j.seedStart = totalSeeds;
			const uint32_t L = rs.length[base + i];
			j.extendSeeds = opt.seedExtendDensity == -1 ? seeds[i].size() : (uint64_t)(opt.seedExtendDensity * L + 1);
			totalSeeds += seeds[i].size();
			jobs.push_back(j);
			jobRead.push_back(i);
		}
		std::vector<gad::DevSeed> devSeeds(totalSeeds);
		parallelFor(jobs.size(), opt.hostThreads, [&](size_t k) {
			const uint32_t i = jobRead[k];
			gad::DevSeed* d = devSeeds.data() + jobs[k].seedStart;
			for (size_t q = 0; q < seeds[i].size(); q++)
			{
				const SeedHit& s = seeds[i][q];
				const int forwardId = s.reverse ? s.nodeID * 2 + 1 : s.nodeID * 2;
				const int backwardId = s.reverse ? s.nodeID * 2 : s.nodeID * 2 + 1;
				gad::DevSeed& e = d[q];
				e.seedGoodness = s.seedGoodness;
				e.clusterSize = s.seedClusterSize;
				e.seqPos = (uint32_t)s.seqPos;
				e.nodeID = s.nodeID;
				e.nodeOffset = (uint32_t)s.nodeOffset;
				e.reverse = s.reverse ? 1 : 0;
				// getInitialSliceExactPosition's split node for both directions
				const uint32_t fn = (uint32_t)s.alignmentGraphNodeId;
				e.fwdNode = fn;
				e.fwdOff = (uint32_t)(s.nodeOffset - g.nodeOff[fn]);
				const auto rpos = g.reversePosition(forwardId, s.nodeOffset);
				const uint32_t bnode = g.unitigNode(backwardId, rpos.second);
				e.bwdNode = bnode;
				e.bwdOff = (uint32_t)(rpos.second - g.nodeOff[bnode]);
			}
		});
		stats.hostMs += msSince(th);
		ReadLoopOutput rlo;
		auto te = std::chrono::steady_clock::now();
		if (opt.batchedDeviceLoop) ctx.runReadLoopsBatched(g, rp, devSeeds, jobs, rlo);
		else ctx.runReadLoops(g, rp, devSeeds, jobs, rlo);
		stats.workers = std::max<uint64_t>(stats.workers, rlo.workers);
		stats.tuneMs += rlo.tuneMs;
		stats.extendCallMs += msSince(te);
		stats.rounds += rlo.launches;
		stats.launches += rlo.launches;
		stats.extensions += rlo.extensions;
		stats.extendKernelMs += rlo.kernelMs;
		stats.wordSteps += rlo.wordSteps;
		stats.reruns += rlo.reruns;
		th = std::chrono::steady_clock::now();
		// alignments in creation order, the sloppy re-sort after each one, then
		// selection, sorts and GAF (src/GraphAligner.h:141-177, src/Aligner.cpp:502-606)
		parallelFor(jobs.size(), opt.hostThreads, [&](size_t k) {
			const uint32_t i = jobRead[k];
			ReadOutcome& o = out[batchStart + i];
			const gad::ReadResult& r = rlo.results[k];
			o.seedsExtended = r.seedsExtended;
			if (r.status == gad::READLOOP_ASSERT)
			{
				o.status = READ_ASSERT;
				return;
			}
			if (r.status != gad::READLOOP_DONE)
			{
				o.status = READ_OVERFLOW;
				return;
			}
			const char* seq = rs.seq(base + i);
			const uint64_t L = rs.length[base + i];
			try
			{
				std::vector<Alignment> al;
				for (uint32_t a = 0; a < r.nAlignments; a++)
				{
					const gad::DevAlignment& da = rlo.alns[rlo.alnOffset[k] + a];
					const uint64_t* raw = rlo.pool.data() + rlo.poolOffset[k] + da.poolOffset;
					const SeedHit& s = seeds[i][da.seedIndex];
					Alignment item = buildAlignment(g, opt, seq, L, s, PackedReader{raw}, da.nb, PackedReader{raw + da.nb}, da.nf);
					if (item.alignmentFailed()) continue;
					item.seedGoodness = s.seedGoodness;
					item.seedIndex = da.seedIndex;
					al.push_back(std::move(item));
					if (opt.sloppyOptimizations) std::sort(al.begin(), al.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentStart < r.alignmentStart; });
				}
				std::vector<Alignment> sel;
				if (!al.empty()) sel = selectAlignments(al, opt.minAlignmentScore, opt.multimapScoreFraction);
				if (opt.records && !clusterOf[i].empty()) opt.records->add(batchStart + i, features[i], (uint32_t)(features[i].size() / kClusterFeatures), clusterOf[i], al, sel);
				if (sel.empty())
				{
					o.status = READ_NO_ALIGNMENT;
					return;
				}
				std::sort(sel.begin(), sel.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentStart < r.alignmentStart; });
				std::sort(sel.begin(), sel.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentXScore > r.alignmentXScore; });
				o.gaf.clear();
				for (const auto& a : sel)
				{
					appendGafLine(o.gaf, g, names[batchStart + i], L, a, opt.cigarMatchMismatchMerge);
					o.gaf += '\n';
				}
				o.status = READ_ALIGNED;
			}
			catch (const ReadAssert&)
			{
				o.status = READ_ASSERT;
				o.gaf.clear();
			}
		});
		stats.hostMs += msSince(th);
	}
	stats.totalMs = msSince(t0);
}

namespace {

// Position of a read's seed loop: the next seed to look at and the counters the
// loop's stop rule reads.
struct LoopCursor
{
	size_t next;
	size_t pending;
	uint64_t worstExtendedSeedScore;
	uint64_t seedsExtended;
};

LoopCursor cursorOf(const ReadState& st)
{
	return LoopCursor{st.next, st.pending, st.worstExtendedSeedScore, st.seedsExtended};
}

// Picks the next seed of the read's loop to extend (src/GraphAligner.h:105-140)
// against the read's alignments so far, moving the cursor; returns false when
// the loop is over.
bool pickNext(const ReadState& st, LoopCursor& c, const AlignOptions& opt)
{
	while (c.next < st.seeds.size())
	{
		const SeedHit& s = st.seeds[c.next];
		if (opt.sloppyOptimizations && s.seedGoodness <= st.seedScoreForEndToEndAln) return false;
		if (c.seedsExtended >= st.extendSeeds && s.seedGoodness < c.worstExtendedSeedScore) return false;
		if (s.seedClusterSize < opt.minSeedClusterSize)
		{
			c.next++;
			continue;
		}
		bool covered = false;
		if (opt.sloppyOptimizations)
		{
			for (const auto& a : st.alignments)
			{
				if (a.alignmentStart <= s.seqPos && a.alignmentEnd >= s.seqPos && a.seedGoodness > s.seedGoodness)
				{
					covered = true;
					break;
				}
			}
		}
		if (!covered)
		{
			for (const auto& a : st.alignments)
			{
				if (exactAlignmentPart(a, s))
				{
					covered = true;
					break;
				}
			}
		}
		if (covered)
		{
			c.next++;
			continue;
		}
		// getTwoDirectionalTrace's check of the seed's read position
		// (src/GraphAligner.h:477-478); only a caller's seed can fail it
		check(s.seqPos < st.readLen);
		c.worstExtendedSeedScore = s.seedGoodness;
		c.seedsExtended++;
		c.pending = c.next;
		return true;
	}
	return false;
}

bool advance(ReadState& st, const AlignOptions& opt)
{
	LoopCursor c = cursorOf(st);
	bool more = pickNext(st, c, opt);
	st.next = c.next;
	st.pending = c.pending;
	st.worstExtendedSeedScore = c.worstExtendedSeedScore;
	st.seedsExtended = c.seedsExtended;
	return more;
}

// The sloppy loop's end-to-end rule after an alignment is added
// (src/GraphAligner.h:160-176): when alignments chained by overlap span the
// whole read, later seeds no better than the chain's worst seed are not tried.
void updateEndToEnd(ReadState& st, uint64_t readLen)
{
	auto& al = st.alignments;
	std::sort(al.begin(), al.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentStart < r.alignmentStart; });
	if (al[0].alignmentStart != 0) return;
	uint64_t minGoodness = al[0].seedGoodness;
	uint64_t contiguousEnd = al[0].alignmentEnd;
	for (size_t k = 1; k < al.size(); k++)
	{
		if (al[k].alignmentStart <= contiguousEnd)
		{
			minGoodness = std::min(minGoodness, al[k].seedGoodness);
			contiguousEnd = std::max(contiguousEnd, al[k].alignmentEnd);
		}
	}
	if (contiguousEnd == readLen) st.seedScoreForEndToEndAln = minGoodness;
}

// Seeds extended together in a read's round r (0-based): the first round only
// the picked seed; later rounds four times as many as the round before, or more when
// few reads remain (up to jobBudget extension jobs in the launch, two per seed:
// a launch lasts as long as its longest extension, so jobs that fill idle
// wavefront slots cost little), never more than maxSeeds.
uint32_t speculationWidth(uint32_t round, uint32_t maxSeeds, size_t activeReads, uint32_t jobBudget)
{
	if (round == 0 || maxSeeds <= 1) return 1;
	uint64_t w = (uint64_t)1 << std::min<uint32_t>(2 * round, 20);
	if (activeReads > 0) w = std::max<uint64_t>(w, jobBudget / (2 * (uint64_t)activeReads));
	return (uint32_t)std::min<uint64_t>(w, maxSeeds);
}

} // namespace

// The host-driven rounds over resident reads [first, first + count). With
// callerSeeds (AlignOneWay's form) read first + i takes callerSeeds[i] as its
// ordered seeds instead of seeding, and its alignments go to rawOut[i] instead
// of selection and GAF text.
void runRounds(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex* idxp, const AlignOptions& opt, const char* const* names, uint32_t first, uint32_t count, ReadOutcome* out, AlignStats& stats, const std::vector<SeedHit>* callerSeeds, std::vector<Alignment>* rawOut)
{
	auto t0 = std::chrono::steady_clock::now();
	// built with -DGA_DEBUG_LOG, GA_ROUND_LOG=1 prints one line per round and
	// group to stderr; the default build has no logging
#ifdef GA_DEBUG_LOG
	const bool roundLog = std::getenv("GA_ROUND_LOG") != nullptr;
#else
	constexpr bool roundLog = false;
#endif
	const ReadSet& rs = ctx.reads();
	if ((uint64_t)first + count > rs.size()) throw std::runtime_error("read range outside the resident reads");
	stats = AlignStats{};
	for (uint32_t i = 0; i < count; i++)
	{
		resetOutcome(out[i]);
		if (rs.length[first + i] > (1u << 24)) throw std::runtime_error("reads longer than 16 Mbp are not supported");
	}
	for (uint32_t batchStart = 0; batchStart < count; batchStart += opt.readsPerBatch)
	{
		const uint32_t bn = std::min<uint32_t>(count - batchStart, opt.readsPerBatch);
		const uint32_t base = first + batchStart;
		std::vector<const char*> bseqs(bn);
		std::vector<uint32_t> blens(bn);
		std::vector<uint64_t> boff(bn);
		for (uint32_t i = 0; i < bn; i++)
		{
			bseqs[i] = rs.seq(base + i);
			blens[i] = rs.length[base + i];
			boff[i] = rs.offset[base + i];
		}
		SeedMatches rb;
		auto tl = std::chrono::steady_clock::now();
		if (callerSeeds) ctx.prepareReads(base, bn);
		else ctx.lookupSeeds(*idxp, base, bn, rb);
		stats.lookupCallMs += msSince(tl);
		stats.lookupMs += rb.lookupMs;
		std::vector<ReadState> states(bn);
		auto th = std::chrono::steady_clock::now();
		// expansion and OrderSeeds of the whole batch on the backend
		std::vector<std::vector<SeedHit>> backendSeeds;
		std::vector<uint8_t> backendAsserted;
		if (opt.deviceSeedOrder && !callerSeeds) ctx.orderedSeeds(g, *idxp, base, bn, rb, opt.minimizerSeedDensity, backendSeeds, backendAsserted);
		// seeds of read i: expansion, ordering and the letter check
		auto seedRead = [&](size_t i) {
			ReadState& st = states[i];
			st.read = (uint32_t)i;
			st.readLen = blens[i];
			ReadOutcome& o = out[batchStart + i];
			try
			{
				if (callerSeeds)
				{
					// AlignOneWay asserts that it has seeds (src/GraphAligner.h:99)
					st.seeds = callerSeeds[batchStart + i];
					check(!st.seeds.empty());
				}
				else if (opt.deviceSeedOrder)
				{
					check(!backendAsserted[i]);
					st.seeds = std::move(backendSeeds[i]);
					if (st.seeds.empty())
					{
						o.status = READ_NO_SEEDS;
						st.done = true;
						return;
					}
				}
				else
				{
					st.seeds = seedsFromMatches(g, *idxp, rb.data() + rb.matchOffset[i], rb.matchCount[i], blens[i], opt.minimizerSeedDensity);
					if (st.seeds.empty())
					{
						o.status = READ_NO_SEEDS;
						st.done = true;
						return;
					}
					orderSeedsChecked(g, st.seeds);
				}
				// ReverseComplement of the read asserts on letters it does not know
				check(allComplementable(bseqs[i], blens[i]));
				st.extendSeeds = opt.seedExtendDensity == -1 ? st.seeds.size() : (uint64_t)(opt.seedExtendDensity * blens[i] + 1);
			}
			catch (const ReadAssert&)
			{
				st.status = READ_ASSERT;
				st.done = true;
			}
		};
		// selection and GAF of one read (src/Aligner.cpp:502-606); its alignments
		// are released afterwards
		auto selectAndWrite = [&](size_t i) {
			ReadState& st = states[i];
			ReadOutcome& o = out[batchStart + i];
			if (o.status == READ_NO_SEEDS) return;
			o.seedsExtended = st.seedsExtended;
			if (rawOut)
			{
				o.status = st.status;
				if (st.status == READ_ALIGNED) rawOut[batchStart + i] = std::move(st.alignments);
				return;
			}
			if (st.status != READ_ALIGNED)
			{
				o.status = st.status;
				return;
			}
			std::vector<Alignment> sel;
			try
			{
				if (!st.alignments.empty()) sel = selectAlignments(st.alignments, opt.minAlignmentScore, opt.multimapScoreFraction);
				if (opt.records && st.nclusters) opt.records->add(batchStart + i, st.features, st.nclusters, st.clusterOf, st.alignments, sel);
				if (sel.empty())
				{
					o.status = READ_NO_ALIGNMENT;
					return;
				}
				std::sort(sel.begin(), sel.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentStart < r.alignmentStart; });
				std::sort(sel.begin(), sel.end(), [](const Alignment& l, const Alignment& r) { return l.alignmentXScore > r.alignmentXScore; });
				o.gaf.clear();
				for (const auto& a : sel)
				{
					appendGafLine(o.gaf, g, names[batchStart + i], blens[i], a, opt.cigarMatchMismatchMerge);
					o.gaf += '\n';
				}
				o.status = READ_ALIGNED;
			}
			catch (const ReadAssert&)
			{
				o.status = READ_ASSERT;
				o.gaf.clear();
			}
		};
		auto finishRead = [&](size_t i) {
			selectAndWrite(i);
			states[i].finished = true;
			std::vector<Alignment>().swap(states[i].alignments);
		};
		// the seed loops of G groups of reads (G = the backend's extension lanes, at
		// most opt.pipelineGroups) run in G host threads, each on its own lane,
		// so one group's host replay overlaps the other groups' kernels
		// Each group seeds its own reads when it starts; the other groups wait
		// until group 0 has seeded, so that their seeding runs while group 0's
		// first launch has the GPU instead of before it. A read without seeds
		// ends in its group's first round.
		std::vector<uint32_t> order(bn);
		for (uint32_t i = 0; i < bn; i++) order[i] = i;
		std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return blens[a] > blens[b]; });
		const uint32_t groups = std::max<uint32_t>(1, std::min<uint32_t>({ctx.extendLanes(), opt.pipelineGroups, (uint32_t)std::max<size_t>(1, order.size())}));
		std::vector<std::vector<uint32_t>> groupReads(groups);
		if (opt.groupByLength)
		{
			// contiguous length classes with equal shares of the bases: the longest
			// reads' group has few reads and launches as long as its longest
			// extension, the other many reads and launches bound by their total
			// work; with one launch at a time the groups fall out of step by
			// themselves (measured on MI355X: 0.84 s per step against 0.91 s for
			// reads dealt round-robin; shares of 35 % or 65 % for the long reads'
			// group were slower)
			uint64_t total = 0, acc = 0;
			for (uint32_t r : order) total += blens[r];
			for (uint32_t r : order)
			{
				const uint32_t grp = (uint32_t)std::min<uint64_t>(groups - 1, acc * groups / std::max<uint64_t>(1, total));
				groupReads[grp].push_back(r);
				acc += blens[r];
			}
		}
		else
			for (size_t k = 0; k < order.size(); k++) groupReads[k % groups].push_back(order[k]);
		for (auto& gr : groupReads) std::sort(gr.begin(), gr.end());
		std::vector<AlignStats> groupStats(groups);
		std::vector<std::exception_ptr> groupErrors(groups);
		const int threads = opt.hostThreads;
		// the other groups' first launches wait until group 0's has returned: the
		// groups start out of phase, so while one group's host work runs another's
		// kernel has the GPU
		std::mutex phaseMutex;
		std::mutex gpuMutex;
		std::condition_variable phaseCv;
		uint32_t firstLaunchesDone = 0;
		struct FirstLaunch
		{
			std::mutex& m;
			std::condition_variable& cv;
			uint32_t& done;
			bool pending = true;
			void signal()
			{
				if (!pending) return;
				pending = false;
				{
					std::lock_guard<std::mutex> lock(m);
					done++;
				}
				cv.notify_all();
			}
			~FirstLaunch() { signal(); }
		};
		// runs finishRead over a list of reads on a thread of its own, one list at a time
		struct Finisher
		{
			std::function<void(size_t)> fn;
			int threads;
			double& ms;
			std::thread worker;
			std::vector<uint32_t> reads;
			std::exception_ptr error;
			void join()
			{
				if (worker.joinable()) worker.join();
				if (error) std::rethrow_exception(error);
			}
			void run(std::vector<uint32_t>&& next)
			{
				join();
				if (next.empty()) return;
				reads = std::move(next);
				worker = std::thread([this]() {
					try
					{
						auto t = std::chrono::steady_clock::now();
						parallelFor(reads.size(), threads, [&](size_t k) { fn(reads[k]); });
						ms += msSince(t);
					}
					catch (...)
					{
						error = std::current_exception();
					}
				});
			}
			~Finisher()
			{
				if (worker.joinable()) worker.join();
			}
		};
		bool group0Seeded = false;
		auto signalSeeded = [&]() {
			{
				std::lock_guard<std::mutex> lock(phaseMutex);
				group0Seeded = true;
			}
			phaseCv.notify_all();
		};
		// the learned proposal's pruning of a group's seeds, then the features of
		// the clusters the loops will see (outcome records)
		std::mutex proposalMutex;
		auto seedProposal = [&](const std::vector<uint32_t>& mine) {
			if (opt.proposal)
			{
				std::vector<std::vector<SeedHit>*> sp;
				std::vector<const char*> sq;
				std::vector<uint32_t> ln;
				for (uint32_t r : mine)
					if (!states[r].done)
					{
						sp.push_back(&states[r].seeds);
						sq.push_back(bseqs[r]);
						ln.push_back(blens[r]);
					}
				ProposalStats ps;
				applyProposal(*opt.proposal, g, idxp ? idxp->maxCount : 0, sp.size(), sp.data(), sq.data(), ln.data(), threads, ps);
				std::lock_guard<std::mutex> lock(proposalMutex);
				stats.proposalClusters += ps.clusters;
				stats.prunedClusters += ps.prunedClusters;
				stats.prunedSeeds += ps.prunedSeeds;
				stats.proposalMs += ps.ms;
			}
			if (opt.records)
				parallelFor(mine.size(), threads, [&](size_t k) {
					ReadState& st = states[mine[k]];
					if (st.done || st.seeds.empty()) return;
					st.nclusters = seedClusters(g, st.seeds, st.clusterOf);
					st.features.resize((size_t)st.nclusters * kClusterFeatures);
					clusterFeatures(g, idxp ? idxp->maxCount : 0, bseqs[st.read], blens[st.read], st.seeds, st.clusterOf, st.nclusters, st.nclusters, st.features.data());
				});
		};
		auto runGroup = [&](uint32_t lane) {
			AlignStats& gs = groupStats[lane];
			Finisher finisher{finishRead, threads, gs.gafMs};
			FirstLaunch firstLaunch{phaseMutex, phaseCv, firstLaunchesDone};
			{
				if (lane > 0)
				{
					std::unique_lock<std::mutex> lock(phaseMutex);
					phaseCv.wait(lock, [&]() { return group0Seeded; });
				}
				auto ts = std::chrono::steady_clock::now();
				const std::vector<uint32_t>& mine = groupReads[lane];
				try
				{
					parallelFor(mine.size(), threads, [&](size_t k) { seedRead(mine[k]); });
				}
				catch (...)
				{
					if (lane == 0) signalSeeded();
					throw;
				}
				if (opt.proposal || opt.records)
					try
					{
						seedProposal(mine);
					}
					catch (...)
					{
						if (lane == 0) signalSeeded();
						throw;
					}
				gs.seedingMs += msSince(ts);
				for (uint32_t r : mine) gs.seeds += states[r].seeds.size();
				if (lane == 0) signalSeeded();
			}
			std::vector<uint32_t> active = groupReads[lane];
			while (!active.empty())
			{
				auto th = std::chrono::steady_clock::now();
				// each read's next seed (the one its loop picks now), then, from its
				// second round on, the seeds its loop would pick after it if those
				// alignments covered nothing: extended in the same launch and replayed
				parallelFor(active.size(), threads, [&](size_t a) {
					ReadState& st = states[active[a]];
					st.specs.clear();
					if (st.done) return; // no seeds, or a letter the read's complement rejects
					try
					{
						if (!st.pendingReady && !advance(st, opt))
						{
							st.done = true;
							return;
						}
						st.pendingReady = false;
						st.specs.push_back(Spec{st.pending, -1, -1});
						const uint32_t k = speculationWidth(st.roundsRun, opt.maxSpeculativeSeeds, active.size(), opt.speculationJobBudget);
						if (k > 1)
						{
							LoopCursor c = cursorOf(st);
							c.next = st.pending + 1;
							try
							{
								while (st.specs.size() < k && pickNext(st, c, opt))
								{
									st.specs.push_back(Spec{c.pending, -1, -1});
									c.next = c.pending + 1;
								}
							}
							catch (const ReadAssert&)
							{
								// a check that fires on a seed picked ahead stops the
								// speculation there; the replayed loop decides whether
								// it reaches that seed and fails
							}
						}
						st.roundsRun++;
						// the specs' extension jobs: from each seed backwards over the
						// reverse-complemented read and forwards over the read
						st.jobs.clear();
						const uint64_t L = blens[active[a]];
						const uint32_t r = active[a];
						for (Spec& sp : st.specs)
						{
							const SeedHit& s = st.seeds[sp.seed];
							int forwardId = s.reverse ? s.nodeID * 2 + 1 : s.nodeID * 2;
							int backwardId = s.reverse ? s.nodeID * 2 : s.nodeID * 2 + 1;
							if (s.seqPos > 0)
							{
								auto rp = g.reversePosition(forwardId, s.nodeOffset);
								uint32_t node = g.unitigNode(backwardId, rp.second);
								ExtJob j{nullptr, (uint32_t)s.seqPos, node, (uint32_t)(rp.second - g.nodeOff[node])};
								j.devOffset = ctx.reverseBase() + boff[r] + (L - s.seqPos);
								sp.bwdJob = (int64_t)st.jobs.size();
								st.jobs.push_back(j);
							}
							if (s.seqPos < L - 1)
							{
								uint32_t node = g.unitigNode(forwardId, s.nodeOffset);
								ExtJob j{nullptr, (uint32_t)(L - s.seqPos - 1), node, (uint32_t)(s.nodeOffset - g.nodeOff[node])};
								j.devOffset = boff[r] + s.seqPos + 1;
								sp.fwdJob = (int64_t)st.jobs.size();
								st.jobs.push_back(j);
							}
						}
					}
					catch (const ReadAssert&)
					{
						st.status = READ_ASSERT;
						st.done = true;
					}
				});
				std::vector<ExtJob> jobs;
				std::vector<uint32_t> extending;
				for (uint32_t r : active)
				{
					ReadState& st = states[r];
					if (st.done) continue;
					const int64_t base = (int64_t)jobs.size();
					jobs.insert(jobs.end(), st.jobs.begin(), st.jobs.end());
					for (Spec& sp : st.specs)
					{
						if (sp.bwdJob >= 0) sp.bwdJob += base;
						if (sp.fwdJob >= 0) sp.fwdJob += base;
					}
					extending.push_back(r);
				}
				{ double ms = msSince(th); gs.hostMs += ms; gs.pickMs += ms; }
				if (extending.empty()) break;
				if (firstLaunch.pending && lane > 0 && !opt.groupByLength)
				{
					std::unique_lock<std::mutex> lock(phaseMutex);
					phaseCv.wait(lock, [&]() { return firstLaunchesDone >= 1; });
				}
				ExtBatch eb;
				auto te = std::chrono::steady_clock::now();
				{
					// one group's launch at a time: a launch lasts as long as its
					// longest extension, which two concurrent launches slow down (and
					// whose shortest extensions then wait for wave slots), while the
					// other group's host work overlaps either way
					std::lock_guard<std::mutex> gpu(gpuMutex);
					ctx.extend(g, opt.ext, jobs, false, eb, ctx.readsBase(), lane);
				}
				gs.extendCallMs += msSince(te);
				firstLaunch.signal();
				gs.rounds++;
				gs.launches += eb.launches;
				gs.reruns += eb.reruns;
				gs.extensions += jobs.size();
				gs.extendKernelMs += eb.kernelMs;
				gs.wordSteps += eb.wordSteps;
				th = std::chrono::steady_clock::now();
				std::atomic<uint64_t> unused{0};
				parallelFor(extending.size(), threads, [&](size_t a) {
					ReadState& st = states[extending[a]];
					const uint32_t r = extending[a];
					const char* seq = bseqs[r];
					const uint64_t L = blens[r];
					size_t si = 0;
					try
					{
						// replay of the loop: the seed it picks is always one of this
						// round's (the first, then any later one the alignments so far
						// leave uncovered) or a seed for the next round
						while (true)
						{
							const Spec& sp = st.specs[si];
							const SeedHit s = st.seeds[sp.seed];
							st.next = sp.seed + 1;
							int64_t jobOf[2] = {-1, -1};
							uint32_t counts[2] = {0, 0};
							for (int dir = 0; dir < 2; dir++)
							{
								int64_t j = dir == 0 ? sp.bwdJob : sp.fwdJob;
								if (j < 0) continue;
								const gad::ExtResult& res = eb.results[j];
								if (res.status == gad::EXT_OVERFLOW)
								{
									st.status = READ_OVERFLOW;
									st.done = true;
									return;
								}
								check(res.status != gad::EXT_ASSERT);
								if (res.status != gad::EXT_OK) continue;
								check(res.ntrace > 0);
								jobOf[dir] = j;
								counts[dir] = res.ntrace;
							}
							Alignment item = buildAlignment(g, opt, seq, L, s, CodeReader(g, eb, jobOf[0] < 0 ? 0 : jobOf[0]), counts[0], CodeReader(g, eb, jobOf[1] < 0 ? 0 : jobOf[1]), counts[1]);
							if (!item.alignmentFailed())
							{
								item.seedGoodness = s.seedGoodness;
								item.seedIndex = (uint32_t)sp.seed;
								st.alignments.push_back(std::move(item));
								if (opt.sloppyOptimizations) updateEndToEnd(st, L);
							}
							if (!advance(st, opt))
							{
								unused += st.specs.size() - si - 1;
								st.done = true;
								break;
							}
							size_t k = si + 1;
							while (k < st.specs.size() && st.specs[k].seed < st.pending) k++;
							if (k < st.specs.size() && st.specs[k].seed == st.pending)
							{
								unused += k - si - 1;
								si = k;
								continue;
							}
							unused += st.specs.size() - si - 1;
							st.pendingReady = true;
							break;
						}
					}
					catch (const ReadAssert&)
					{
						st.status = READ_ASSERT;
						st.done = true;
					}
				});
				gs.unusedExtensions += unused.load();
				{
					double ms = msSince(th);
					gs.hostMs += ms;
					gs.replayMs += ms;
					if (roundLog)
					{
						// the launch's heaviest extension (forward word steps), which
						// bounds its duration
						size_t top = 0, last = 0;
						uint64_t c0 = ~0ull;
						for (size_t j = 0; j < jobs.size(); j++)
						{
							if (eb.results[j].wordSteps > eb.results[top].wordSteps) top = j;
							if (eb.results[j].clockEnd > eb.results[last].clockEnd) last = j;
							c0 = std::min<uint64_t>(c0, eb.results[j].clockStart);
						}
						if (!jobs.empty() && eb.results[top].clockEnd)
						{
							// device clock ticks of 10 ns
							auto rel = [&](uint64_t t) { return (double)(t - c0) * 1e-5; };
							std::fprintf(stderr, "[rounds]   clocks (ms from the first start): heaviest %.1f-%.1f, last to end: len %u, %.1f-%.1f\n",
							             rel(eb.results[top].clockStart), rel(eb.results[top].clockEnd), jobs[last].len, rel(eb.results[last].clockStart), rel(eb.results[last].clockEnd));
						}
						std::fprintf(stderr, "[rounds] group %u round %llu: %zu reads, %zu jobs, extend %.1f ms (kernel %.1f ms), replay %.1f ms, ends at %.1f ms; heaviest: len %u, %llu word steps, %llu node calcs, %u slices, band %u, trace %u\n", lane,
						             (unsigned long long)gs.rounds, extending.size(), jobs.size(), msSince(te) - ms, eb.kernelMs, ms, msSince(t0),
						             jobs.empty() ? 0u : jobs[top].len, jobs.empty() ? 0ull : (unsigned long long)eb.results[top].wordSteps,
						             jobs.empty() ? 0ull : (unsigned long long)eb.results[top].nodeCalcs, jobs.empty() ? 0u : eb.results[top].nslices,
						             jobs.empty() ? 0u : eb.results[top].maxBand, jobs.empty() ? 0u : eb.results[top].ntrace);
					}
				}
				std::vector<uint32_t> still, ended;
				for (uint32_t r : active)
					(states[r].done ? ended : still).push_back(r);
				active.swap(still);
				// reads whose loops ended are written out by a helper thread while
				// the group's next round runs
				finisher.run(std::move(ended));
			}
			finisher.join();
		};
		if (groups == 1) runGroup(0);
		else
		{
			std::vector<std::thread> pool;
			for (uint32_t lane = 0; lane < groups; lane++)
				pool.emplace_back([&, lane]() {
					try
					{
						runGroup(lane);
					}
					catch (...)
					{
						groupErrors[lane] = std::current_exception();
					}
				});
			for (auto& t : pool) t.join();
			for (auto& e : groupErrors)
				if (e) std::rethrow_exception(e);
		}
		for (const AlignStats& gs : groupStats)
		{
			stats.seeds += gs.seeds;
			stats.rounds = std::max(stats.rounds, gs.rounds);
			stats.launches += gs.launches;
			stats.reruns += gs.reruns;
			stats.extensions += gs.extensions;
			stats.extendKernelMs += gs.extendKernelMs;
			stats.wordSteps += gs.wordSteps;
			stats.extendCallMs = std::max(stats.extendCallMs, gs.extendCallMs);
			stats.unusedExtensions += gs.unusedExtensions;
		}
		// the groups' host phases overlap: count the longest group's
		{
			double hostMax = 0, pickMax = 0, replayMax = 0, gafMax = 0, seedMax = 0;
			for (const AlignStats& gs : groupStats)
			{
				seedMax = std::max(seedMax, gs.seedingMs);
				hostMax = std::max(hostMax, gs.hostMs);
				pickMax = std::max(pickMax, gs.pickMs);
				replayMax = std::max(replayMax, gs.replayMs);
				gafMax = std::max(gafMax, gs.gafMs);
			}
			stats.seedingMs += seedMax;
			stats.hostMs += hostMax + seedMax;
			stats.pickMs += pickMax;
			stats.replayMs += replayMax;
			stats.gafMs += gafMax;
		}
		// selection and GAF of the reads the seed loops did not reach
		th = std::chrono::steady_clock::now();
		parallelFor(bn, opt.hostThreads, [&](size_t i) {
			if (!states[i].finished) finishRead(i);
		});
		{ double ms = msSince(th); stats.hostMs += ms; stats.gafMs += ms; }
	}
	stats.totalMs = msSince(t0);
	if (roundLog) std::fprintf(stderr, "[rounds] done at %.1f ms\n", stats.totalMs);
}

void alignResident(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, const AlignOptions& opt, const char* const* names, uint32_t first, uint32_t count, ReadOutcome* out, AlignStats& stats)
{
	ctx.beginAlignmentCall();
	if (opt.deviceSeedLoop) alignResidentDeviceLoop(ctx, g, idx, opt, names, first, count, out, stats);
	else runRounds(ctx, g, &idx, opt, names, first, count, out, stats, nullptr, nullptr);
}

void alignOneWayResident(ExtendBackend& ctx, const GraphHost& g, const AlignOptions& opt, uint32_t first, uint32_t count, const std::vector<SeedHit>* seeds, std::vector<Alignment>* alignments, ReadOutcome* out, AlignStats& stats)
{
	for (uint32_t i = 0; i < count; i++) alignments[i].clear();
	ctx.beginAlignmentCall();
	runRounds(ctx, g, nullptr, opt, nullptr, first, count, out, stats, seeds, alignments);
}

void alignReads(ExtendBackend& ctx, const GraphHost& g, const MinimizerIndex& idx, const AlignOptions& opt, const std::vector<const char*>& names, const std::vector<const char*>& seqs, const std::vector<uint32_t>& lens, std::vector<ReadOutcome>& out, AlignStats& stats)
{
	auto t0 = std::chrono::steady_clock::now();
	for (uint32_t l : lens)
		if (l > (1u << 24)) throw std::runtime_error("reads longer than 16 Mbp are not supported");
	ctx.uploadReads(seqs, lens);
	out.resize(seqs.size());
	alignResident(ctx, g, idx, opt, names.data(), 0, (uint32_t)seqs.size(), out.data(), stats);
	stats.totalMs = msSince(t0);
}

} // namespace ga
