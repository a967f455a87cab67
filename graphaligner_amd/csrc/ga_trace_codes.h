// Compact form of an extension's trace for the copy from HBM to the host,
// shared by the gfx950 encoder kernel and the host decoder (same source, so the
// CPU suite checks both against the oracle).
//
// A trace is a list of packed items (split node, offset, read position,
// node-switch flag; see packTrace) written end first, so each item lies at or
// before the one written before it: on the same read row or one row up, in the
// same column, one column to the left, or in the last column of an in-neighbour.
// Item k > 0 is stored as one byte relative to item k - 1:
//   bit 0      the item's node-switch flag
//   bit 1      1 when the item is one read row above item k - 1
//   bits 2..7  0: same node and column; 1: same node, one column left;
//              2 + m: last column of the m-th in-neighbour of item k - 1's node
// and the byte 0xFF (escape) stands for an item kept whole in a side list. Item 0
// and anything the byte cannot express are escapes; the encoder also escapes any
// item whose byte would not decode back to the same 64 bits.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GA_TC_HD __host__ __device__ inline
#else
#define GA_TC_HD inline
#endif

namespace gad {

constexpr uint32_t kTraceEscape = 0xFF;
constexpr uint32_t kTraceMaxNeighbour = 61; // kinds 2..62 keep every code below 0xFF

struct TraceGraph
{
	const uint32_t* nodeLen;
	const uint32_t* inOff;
	const uint32_t* inAdj;
};

// item layout of packTrace: node in the low 32 bits; high word
// (seqPos + 1) << 7 | offset << 1 | nodeSwitch
GA_TC_HD uint64_t traceDecode(const TraceGraph& g, uint64_t prev, uint32_t code)
{
	const uint32_t pn = (uint32_t)prev, ph = (uint32_t)(prev >> 32);
	const uint32_t kind = code >> 2;
	uint32_t node = pn, off = (ph >> 1) & 63;
	if (kind == 1) off -= 1;
	else if (kind >= 2)
	{
		node = g.inAdj[g.inOff[pn] + (kind - 2)];
		off = g.nodeLen[node] - 1;
	}
	const uint32_t seq1 = (ph >> 7) - ((code >> 1) & 1);
	const uint32_t hi = (seq1 << 7) | (off << 1) | (code & 1);
	return ((uint64_t)hi << 32) | node;
}

GA_TC_HD uint32_t traceCode(const TraceGraph& g, uint64_t prev, uint64_t cur)
{
	const uint32_t pn = (uint32_t)prev, cn = (uint32_t)cur;
	const uint32_t ph = (uint32_t)(prev >> 32), ch = (uint32_t)(cur >> 32);
	const uint32_t up = (ph >> 7) - (ch >> 7);
	if (up > 1) return kTraceEscape;
	const uint32_t poff = (ph >> 1) & 63, coff = (ch >> 1) & 63;
	uint32_t kind = kTraceEscape;
	if (cn == pn && coff == poff) kind = 0;
	else if (cn == pn && coff + 1 == poff) kind = 1;
	else
	{
		const uint32_t b = g.inOff[pn], e = g.inOff[pn + 1];
		for (uint32_t k = b; k < e && k - b < kTraceMaxNeighbour; k++)
		{
			if (g.inAdj[k] == cn)
			{
				kind = 2 + (k - b);
				break;
			}
		}
		if (kind == kTraceEscape) return kTraceEscape;
	}
	const uint32_t code = (kind << 2) | (up << 1) | (ch & 1);
	return traceDecode(g, prev, code) == cur ? code : kTraceEscape;
}

} // namespace gad
