// One read's seed loop on the device: GraphAligner::AlignOneWay
// (src/GraphAligner.h:94-181) with getAlignmentFromSeed's trace checks
// (:475-520, :522-601, :631-692) and exactAlignmentPart (:403-457). Each
// extension runs through runRequest; alignments are recorded in creation order
// with their raw traces, and the host later rebuilds the alignment items,
// replays the reference's sorts, selects and writes GAF.
//
// Only the loop's decisions are made here: which seeds are extended, the
// alignments' read intervals, and the assertions that drop a read. The trace
// score recount and the character checks happen on the host; a read whose
// characters would assert there is rejected before it reaches this loop.
#pragma once
#include "ga_extend_core.h"

namespace gad {

struct RawItem
{
	uint32_t node;
	uint32_t off;
	uint64_t seqPos; // ~0 for row -1
	bool sw;
};

GA_HD RawItem unpackTrace(uint64_t v)
{
	RawItem r;
	uint32_t hi = (uint32_t)(v >> 32);
	r.node = (uint32_t)v;
	r.off = (hi >> 1) & 63;
	r.seqPos = (uint64_t)((int64_t)(hi >> 7) - 1);
	r.sw = (hi & 1) != 0;
	return r;
}

// Bigraph coordinates of a merged-trace item.
struct MergedPos
{
	uint64_t node;
	uint64_t off;
	uint64_t seqPos;
};

// The merged trace of getAlignmentFromSeed as a view over the two raw traces:
// the fixed backward part (fixReverseTraceSeqPosAndOrder keeps raw order, so
// item m is raw backward item m in reverse orientation) without its last item
// when a forward part follows, then the fixed forward part (raw forward items
// in reverse order).
struct MergedView
{
	const uint64_t* bwd;
	const uint64_t* fwd;
	uint32_t nbUsed;
	uint32_t nf;
	uint64_t seedPos;
	GA_HD uint32_t size() const { return nbUsed + nf; }
};

GA_HD MergedView mergedView(const uint64_t* bwd, uint32_t nb, const uint64_t* fwd, uint32_t nf, uint64_t seedPos)
{
	MergedView v;
	v.bwd = bwd;
	v.fwd = fwd;
	v.nbUsed = (nb > 0 && nf > 0) ? nb - 1 : nb;
	v.nf = nf;
	v.seedPos = seedPos;
	return v;
}

// src/AlignmentGraph.cpp:830-847 applied to a split-node position
GA_HD void reversedBigraph(const DevGraph& g, uint32_t node, uint32_t off, uint64_t& bigNode, uint64_t& bigOff)
{
	int32_t id = g.nodeId[node];
	int32_t rev = (id % 2 == 0) ? (id / 2) * 2 + 1 : (id / 2) * 2;
	bigNode = (uint64_t)(int64_t)rev;
	bigOff = (uint64_t)g.bigSize[node] - ((uint64_t)g.nodeOff[node] + off) - 1;
}

GA_HD MergedPos mergedAt(const DevGraph& g, const MergedView& v, uint32_t m)
{
	MergedPos p;
	if (m < v.nbUsed)
	{
		RawItem r = unpackTrace(v.bwd[m]);
		p.seqPos = (v.seedPos - 1) - r.seqPos;
		reversedBigraph(g, r.node, r.off, p.node, p.off);
	}
	else
	{
		RawItem q = unpackTrace(v.fwd[v.nf - 1 - (m - v.nbUsed)]);
		p.seqPos = q.seqPos + v.seedPos + 1;
		p.node = (uint64_t)(int64_t)g.nodeId[q.node];
		p.off = (uint64_t)g.nodeOff[q.node] + q.off;
	}
	return p;
}

struct ReadLoop
{
	DevGraph g;
	ReadLoopParams p;
	const char* reads;
	uint64_t* pool;
	DevAlignment* alns;
	uint32_t nAln;
	uint32_t alnCap;
	uint32_t poolUsed;
	uint32_t poolCap;
	int32_t status;
	uint32_t overflowWhat;

	GA_HD void fail(int32_t st, uint32_t what = 0)
	{
		if (status == READLOOP_DONE)
		{
			status = st;
			overflowWhat = what;
		}
	}
	GA_HD void check(bool c)
	{
		if (!c) fail(READLOOP_ASSERT);
	}

	GA_HD MergedView viewOf(const DevAlignment& a, uint64_t seedPos) const
	{
		return mergedView(pool + a.poolOffset, a.nb, pool + a.poolOffset + a.nb, a.nf, seedPos);
	}

	// GraphAligner::exactAlignmentPart (src/GraphAligner.h:403-457)
	GA_HD bool exactAlignmentPart(const MergedView& v, const DevSeed& s)
	{
		const uint32_t n = v.size();
		const uint64_t first = mergedAt(g, v, 0).seqPos;
		const uint64_t last = mergedAt(g, v, n - 1).seqPos;
		check(last > first);
		if (status != READLOOP_DONE) return false;
		if (last < s.seqPos) return false;
		if (first > s.seqPos) return false;
		uint64_t high = n, low = 0;
		uint64_t mid = ((uint64_t)s.seqPos - first) / (last - first);
		// every pass narrows [low, high), so n + 1 passes always suffice
		for (uint32_t guard = 0; guard <= n; guard++)
		{
			if (mid >= n)
			{
				fail(READLOOP_ASSERT);
				return false;
			}
			uint64_t sp = mergedAt(g, v, (uint32_t)mid).seqPos;
			if (sp == s.seqPos) break;
			if (sp < s.seqPos)
			{
				low = mid;
				mid = (high + low) / 2;
				if (mid == low) mid += 1;
				if (mid >= n)
				{
					fail(READLOOP_ASSERT);
					return false;
				}
				sp = mergedAt(g, v, (uint32_t)mid).seqPos;
			}
			if (sp > s.seqPos)
			{
				high = mid;
				mid = (high + low) / 2;
			}
			if (!(low < mid) || !(mid < high))
			{
				fail(READLOOP_ASSERT);
				return false;
			}
		}
		const uint64_t compareNode = (uint64_t)(int64_t)(s.nodeID * 2 + (s.reverse ? 1 : 0));
		uint64_t down = mid;
		while (true)
		{
			MergedPos q = mergedAt(g, v, (uint32_t)down);
			if (q.seqPos != s.seqPos) break;
			if (compareNode == q.node && (uint64_t)s.nodeOffset == q.off) return true;
			if (down == 0) break;
			down--;
		}
		uint64_t up = mid;
		while (true)
		{
			MergedPos q = mergedAt(g, v, (uint32_t)up);
			if (q.seqPos != s.seqPos) break;
			if (compareNode == q.node && (uint64_t)s.nodeOffset == q.off) return true;
			up++;
			if (up == n) break;
		}
		return false;
	}

	// GraphAligner::verifyTrace (src/GraphAligner.h:746-766) on a raw trace read
	// start first (raw item count - 1 down to 0)
	GA_HD void verifyRaw(const uint64_t* raw, uint32_t n)
	{
		for (uint32_t k = 1; k < n && status == READLOOP_DONE; k++)
		{
			RawItem o = unpackTrace(raw[n - k]);
			RawItem c = unpackTrace(raw[n - 1 - k]);
			check(c.seqPos != o.seqPos || c.node != o.node || c.off != o.off);
			bool seqStep = c.seqPos == o.seqPos || c.seqPos == o.seqPos + 1;
			if (o.node == c.node)
			{
				bool inside = (c.off == o.off || c.off == o.off + 1) && seqStep;
				bool wrap = o.off == g.nodeLen[c.node] - 1 && c.off == 0 && seqStep;
				check(inside || wrap);
				continue;
			}
			bool edge = false;
			for (uint32_t e = g.outOff[o.node]; e < g.outOff[o.node + 1]; e++)
				if (g.outAdj[e] == c.node) edge = true;
			check(edge);
			check(seqStep);
			check(o.off == g.nodeLen[o.node] - 1);
			check(c.off == 0);
		}
	}

	// sloppy end-to-end check of AlignOneWay (src/GraphAligner.h:160-177): the
	// alignments reachable from one starting at 0 by a sweep in start order
	GA_HD uint64_t endToEndGoodness(uint64_t readLen, uint64_t current) const
	{
		uint64_t minGoodness = 0, contiguousEnd = 0;
		bool any = false, haveLast = false;
		uint64_t lastStart = 0;
		// visit the alignments in increasing start order, one start value at a
		// time (the order among equal starts does not change the result)
		for (uint32_t round = 0; round < nAln; round++)
		{
			uint64_t next = ~0ull;
			for (uint32_t i = 0; i < nAln; i++)
			{
				uint64_t st = alns[i].start;
				if ((!haveLast || st > lastStart) && st < next) next = st;
			}
			if (next == ~0ull) break;
			for (uint32_t i = 0; i < nAln; i++)
			{
				if (alns[i].start != next) continue;
				if (!any)
				{
					if (next != 0) return current;
					any = true;
					minGoodness = alns[i].seedGoodness;
					contiguousEnd = alns[i].end;
				}
				else if (alns[i].start <= contiguousEnd)
				{
					if (alns[i].seedGoodness < minGoodness) minGoodness = alns[i].seedGoodness;
					if (alns[i].end > contiguousEnd) contiguousEnd = alns[i].end;
				}
			}
			lastStart = next;
			haveLast = true;
		}
		if (any && contiguousEnd == readLen) return minGoodness;
		return current;
	}
};

// Runs one read's seed loop. rawTrace holds two regions of job.traceCap items.
// queueStore/queueLds: the extensions' band queue in LDS (as the single-lane
// extension kernel keeps it) when the capacities fit it.
GA_HD void runReadLoop(const DevGraph& g, const ReadLoopParams& p, const ReadJob& job, const char* reads, const DevSeed* seeds, uint64_t* rawTrace, uint64_t* pool, DevAlignment* alns, unsigned char* arena, const ArenaCaps& caps, ReadResult& out, uint64_t* colStore = nullptr, const uint64_t* planes = nullptr, bool colStoreLds = false, uint64_t* queueStore = nullptr, bool queueLds = false)
{
	ReadLoop L;
	L.g = g;
	L.p = p;
	L.reads = reads;
	L.pool = pool;
	L.alns = alns;
	L.nAln = 0;
	L.alnCap = job.alnCap;
	L.poolUsed = 0;
	L.poolCap = job.poolCap;
	L.status = READLOOP_DONE;
	L.overflowWhat = 0;
	out.wordSteps = 0;
	out.extensions = 0;
	out.seedsExtended = 0;
	const uint64_t len = job.len;
	uint64_t e2e = 0, worst = 0, extended = 0;
	uint64_t* bwdRaw = rawTrace;
	uint64_t* fwdRaw = rawTrace + job.traceCap;
	for (uint32_t i = 0; i < job.seedCount && L.status == READLOOP_DONE; i++)
	{
		const DevSeed s = seeds[i];
		if (p.sloppy && s.seedGoodness <= e2e) break;
		if (extended >= job.extendSeeds && s.seedGoodness < worst) break;
		if (s.clusterSize < p.minSeedClusterSize) continue;
		bool covered = false;
		if (p.sloppy)
		{
			for (uint32_t a = 0; a < L.nAln; a++)
			{
				if ((uint64_t)L.alns[a].start <= s.seqPos && (uint64_t)L.alns[a].end >= s.seqPos && L.alns[a].seedGoodness > s.seedGoodness)
				{
					covered = true;
					break;
				}
			}
		}
		if (!covered)
		{
			for (uint32_t a = 0; a < L.nAln && L.status == READLOOP_DONE; a++)
			{
				const DevAlignment al = L.alns[a];
				if (L.exactAlignmentPart(L.viewOf(al, seeds[al.seedIndex].seqPos), s))
				{
					covered = true;
					break;
				}
			}
		}
		if (L.status != READLOOP_DONE) break;
		if (covered) continue;
		worst = s.seedGoodness;
		extended++;
		// getTwoDirectionalTrace: backward then forward extension
		uint32_t nb = 0, nf = 0;
		bool bwdOk = false, fwdOk = false;
		for (int dir = 0; dir < 2; dir++)
		{
			ExtRequest r;
			if (dir == 0)
			{
				if (s.seqPos == 0) continue;
				r.seqOffset = job.revOffset + (len - s.seqPos);
				r.len = s.seqPos;
				r.startNode = s.bwdNode;
				r.startOffset = s.bwdOff;
				r.traceOffset = 0;
			}
			else
			{
				if ((uint64_t)s.seqPos >= len - 1) continue;
				r.seqOffset = job.fwdOffset + s.seqPos + 1;
				r.len = (uint32_t)(len - s.seqPos - 1);
				r.startNode = s.fwdNode;
				r.startOffset = s.fwdOff;
				r.traceOffset = 0;
			}
			r.traceCap = job.traceCap;
			ExtResult res;
			runRequest(g, p.ext, r, reads, dir == 0 ? bwdRaw : fwdRaw, arena, caps, nullptr, res, colStore, planes, colStoreLds, queueStore, queueLds);
			out.wordSteps += res.wordSteps;
			out.extensions++;
			if (res.status == EXT_OVERFLOW)
			{
				L.fail(READLOOP_OVERFLOW, 16 + res.overflowWhat);
				break;
			}
			if (res.status == EXT_ASSERT || (res.status == EXT_OK && res.ntrace == 0))
			{
				L.fail(READLOOP_ASSERT);
				break;
			}
			if (res.status != EXT_OK) continue;
			// the traceback ends in row -1 of the seed column
			const uint64_t* raw = dir == 0 ? bwdRaw : fwdRaw;
			RawItem startItem = unpackTrace(raw[res.ntrace - 1]);
			if (dir == 0)
			{
				L.check(startItem.seqPos == ~0ull && startItem.node == s.bwdNode && startItem.off == s.bwdOff);
				nb = res.ntrace;
				bwdOk = true;
			}
			else
			{
				L.check(startItem.seqPos == ~0ull && startItem.node == s.fwdNode && startItem.off == s.fwdOff);
				nf = res.ntrace;
				fwdOk = true;
			}
		}
		if (L.status != READLOOP_DONE) break;
		if (fwdOk) L.verifyRaw(fwdRaw, nf);
		if (bwdOk) L.verifyRaw(bwdRaw, nb);
		if (L.status != READLOOP_DONE) break;
		// fixReverseTraceSeqPosAndOrder / fixForwardTraceSeqPos position checks
		for (uint32_t k = 0; k < nb; k++) L.check(((uint64_t)s.seqPos - 1) - unpackTrace(bwdRaw[k]).seqPos < len);
		for (uint32_t k = 0; k < nf; k++) L.check(unpackTrace(fwdRaw[k]).seqPos + (uint64_t)s.seqPos + 1 < len);
		if (L.status != READLOOP_DONE) break;
		if (!bwdOk && !fwdOk) continue;
		MergedView v = mergedView(bwdRaw, nb, fwdRaw, nf, s.seqPos);
		if (bwdOk && fwdOk)
		{
			// the backward part's last item equals the forward part's first
			RawItem rl = unpackTrace(bwdRaw[nb - 1]);
			uint64_t bn, bo;
			reversedBigraph(g, rl.node, rl.off, bn, bo);
			uint64_t bs = ((uint64_t)s.seqPos - 1) - rl.seqPos;
			MergedPos f0 = mergedAt(g, v, v.nbUsed);
			L.check(bn == f0.node && bo == f0.off && bs == f0.seqPos);
			if (L.status != READLOOP_DONE) break;
		}
		const uint32_t n = v.size();
		L.check(n > 0);
		if (L.status != READLOOP_DONE) break;
		const uint64_t start = mergedAt(g, v, 0).seqPos;
		const uint64_t endPos = mergedAt(g, v, n - 1).seqPos;
		L.check(endPos < len);
		if (L.status != READLOOP_DONE) break;
		if (endPos + 1 == start) continue; // alignmentFailed
		if (L.nAln >= L.alnCap)
		{
			L.fail(READLOOP_OVERFLOW, 1);
			break;
		}
		if ((uint64_t)L.poolUsed + nb + nf > L.poolCap)
		{
			L.fail(READLOOP_OVERFLOW, 2);
			break;
		}
		DevAlignment a;
		a.seedGoodness = s.seedGoodness;
		a.poolOffset = L.poolUsed;
		a.seedIndex = i;
		a.start = (uint32_t)start;
		a.end = (uint32_t)(endPos + 1);
		a.nb = nb;
		a.nf = nf;
		for (uint32_t k = 0; k < nb; k++) pool[L.poolUsed + k] = bwdRaw[k];
		for (uint32_t k = 0; k < nf; k++) pool[L.poolUsed + nb + k] = fwdRaw[k];
		L.poolUsed += nb + nf;
		alns[L.nAln++] = a;
		if (p.sloppy) e2e = L.endToEndGoodness(len, e2e);
	}
	out.status = L.status;
	out.overflowWhat = L.overflowWhat;
	out.nAlignments = L.nAln;
	out.poolUsed = L.poolUsed;
	out.seedsExtended = (uint32_t)extended;
}

} // namespace gad
