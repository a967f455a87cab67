// Seed extension on one wavefront: the Myers bit-parallel banded DP over the
// alignment graph (64-row slices, component-ordered band queue, X-drop) and the
// traceback, for one extension request.
//
// All 64 lanes of the wave run the same control flow on the same (uniform)
// values; the 64-bit VP/VN column words are uniform registers. Work over the 64
// rows of a column (cellwise minimum of two columns, per-row X-scores, row
// comparisons) and over band/queue entries (node search, queue minimum) is
// spread over the lanes and finished with ballots and butterfly reductions.
//
// The logic restates src/GraphAlignerBitvectorCommon.h and
// src/GraphAlignerBitvectorBanded.h (getReverseTraceFromSeed, single-seed path):
//  * calculateNodeInner          -> calcNode
//  * calculateSlice/getXdropSlices -> forward
//  * getReverseTraceFromTable(ExactEndPos) and pickBacktrace* -> traceback
// with the band queue's ties on (component, score) broken by insertion order.
//
// The same source builds for gfx950 and, with GA_WAVE_EMULATION defined, as a
// host library whose lane helpers loop over the 64 lanes; that build exists only
// for the CPU test suite.
#pragma once
#include <stdint.h>
#include "ga_device_types.h"

#if defined(GA_WAVE_EMULATION)
#include <algorithm>
#define GA_HD inline
#define GA_HHD inline
#elif defined(GA_SIMT_EMULATION)
// host test build running the lane code on 64 threads; the harness defines the
// wave intrinsics before including this header
#define GA_HD inline
#define GA_HHD inline
#elif defined(GA_THREAD_PER_EXTENSION) && defined(GA_THREAD_HOST_TEST)
// host test build of the one-extension-per-thread code
#include <algorithm>
#define GA_HD inline
#define GA_HHD inline
#elif defined(GA_THREAD_PER_EXTENSION)
// gfx950 build with one extension per thread: no cross-lane operations. Every
// routine is inlined: an out-of-line call would take the extension's state by
// pointer into scratch memory and its arguments in vector registers, losing the
// single-lane kernel's uniform (scalar) values and raising the register count
#include <hip/hip_runtime.h>
#define GA_HD __device__ __forceinline__
#define GA_HHD __host__ __device__ inline
#else
#include <hip/hip_runtime.h>
#define GA_HD __device__ __forceinline__
#define GA_HHD __host__ __device__ inline
#endif

#ifndef GA_COLD
#define GA_COLD GA_HD
#endif

namespace gad {

// The gfx950 extension kernels run one lane per workgroup. Reading a value back
// from the first (only) lane places it in scalar registers, so the column loop's
// 64-bit word arithmetic and its branches run on the scalar unit. Identity in
// every other build.
#if defined(GA_THREAD_PER_EXTENSION) && !defined(GA_THREAD_HOST_TEST)
__device__ __forceinline__ uint32_t uni32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int32_t uni32s(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
	return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) | (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}
#else
GA_HD uint32_t uni32(uint32_t v) { return v; }
GA_HD int32_t uni32s(int32_t v) { return v; }
GA_HD uint64_t uni64(uint64_t v) { return v; }
#endif

constexpr int32_t SMAX = 0x7fffffff;
constexpr int32_t SMIN = (int32_t)0x80000000;
constexpr uint64_t ONES = ~(uint64_t)0;
constexpr uint64_t LASTBIT = (uint64_t)1 << 63;
constexpr uint32_t NIL = 0xffffffffu;
constexpr uint32_t SKIPFIRST = 0x80000000u;
constexpr uint32_t LISTEND = 0x7fffffffu; // end of a queued-edge list (low 31 bits of exNext)
constexpr uint32_t kMaxPopsPerSlice = 1u << 20;
// a node's columns for the backtrace: five 64-bit words per column and 64 scores
constexpr uint32_t kColumnWords = 64 * 5 + 32;

struct Slice
{
	uint64_t VP;
	uint64_t VN;
	int32_t end;
};

// Match masks of rows j..j+63 for A, C, G, T
struct Eq4
{
	uint64_t m[4];
};
// mask of the base with 2-bit code (A0 C1 G2 T3), by bit arithmetic: a select
// between array elements may be compiled into an indexed load from memory
GA_HD uint64_t eqSelect(const Eq4& e, uint32_t code)
{
	const uint64_t b0 = (uint64_t)0 - (uint64_t)(code & 1);
	const uint64_t b1 = (uint64_t)0 - (uint64_t)((code >> 1) & 1);
	const uint64_t lo = e.m[0] ^ ((e.m[0] ^ e.m[1]) & b0);
	const uint64_t hi = e.m[2] ^ ((e.m[2] ^ e.m[3]) & b0);
	return lo ^ ((lo ^ hi) & b1);
}

GA_HD int32_t popc(uint64_t x) { return __builtin_popcountll(x); }
GA_HD int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
GA_HD int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
GA_HD Slice mkslice(uint64_t vp, uint64_t vn, int32_t end)
{
	Slice s;
	s.VP = vp;
	s.VN = vn;
	s.end = end;
	return s;
}
GA_HD bool same(const Slice& a, const Slice& b) { return a.VP == b.VP && a.VN == b.VN && a.end == b.end; }
// WordSlice::getScoreBeforeStart (src/WordSlice.h:244-247)
GA_HD int32_t before(const Slice& s) { return s.end - popc(s.VP) + popc(s.VN); }
// WordSlice::getValue (src/WordSlice.h:177-186)
GA_HD int32_t valueAt(const Slice& s, int row)
{
	uint64_t mask = row < 63 ? (ONES << (row + 1)) : 0;
	return s.end + popc(s.VN & mask) - popc(s.VP & mask);
}
// a column's step data for the backtrace: the horizontal deltas of the step
// that produced it (Ph: row r scores one more than in the column before, Mh: one
// less) and its match mask
struct ColStep
{
	uint64_t ph, mh, eq;
};
GA_HD ColStep mkstep(uint64_t ph, uint64_t mh, uint64_t eq)
{
	ColStep t;
	t.ph = ph;
	t.mh = mh;
	t.eq = eq;
	return t;
}
// getSourceSliceFromScore (src/GraphAlignerBitvectorCommon.h:1056-1060)
GA_HD Slice sourceSlice(int32_t score) { return mkslice(ONES, 0, score + 64); }

// getNextSlice (src/GraphAlignerBitvectorCommon.h:242-262)
GA_HD Slice nextSlice(uint64_t Eq, Slice s, uint64_t hinP, uint64_t hinN, uint64_t& Ph, uint64_t& Mh)
{
	uint64_t Xv = Eq | s.VN;
	Eq |= hinN;
	uint64_t Xh = (((Eq & s.VP) + s.VP) ^ s.VP) | Eq;
	Ph = s.VN | ~(Xh | s.VP);
	Mh = s.VP & Xh;
	uint64_t tempMh = (Mh << 1) | hinN;
	uint64_t tempPh = (Ph << 1) | hinP;
	Slice r;
	r.VP = tempMh | ~(Xv | tempPh);
	r.VN = tempPh & Xv;
	r.end = s.end - (int32_t)(Mh >> 63) + (int32_t)(Ph >> 63);
	return r;
}

// Candidate rows of WordSlice::maxXScoreLocalMinima (src/WordSlice.h:313-334)
GA_HD uint64_t maxXCandidates(const Slice& s, int cells, uint64_t extraMask)
{
	uint64_t pc = ~s.VP;
	uint64_t m = s.VP & (pc - s.VP);
	m >>= 1;
	m |= LASTBIT & (pc | ~(pc - s.VP)) & ~s.VP;
	m &= extraMask;
	m |= 1;
	if (cells < 64)
	{
		m |= (uint64_t)1 << (cells - 1);
		m &= ~(ONES << cells);
	}
	else
	{
		m |= LASTBIT;
	}
	return m;
}

// 4-bit base set of a read character (bit0 A, bit1 C, bit2 G, bit3 T), 16 for a
// character that makes the reference's characterMatch assert (src/GraphAlignerCommon.h:273-276).
GA_HHD uint32_t readBaseSet(unsigned char c)
{
	switch (c)
	{
		case 'A': case 'a': return 1;
		case 'C': case 'c': return 2;
		case 'G': case 'g': return 4;
		case 'T': case 't': case 'U': case 'u': return 8;
		case 'R': case 'r': return 5;
		case 'Y': case 'y': return 10;
		case 'K': case 'k': return 12;
		case 'M': case 'm': return 3;
		case 'S': case 's': return 6;
		case 'W': case 'w': return 9;
		case 'B': case 'b': return 14;
		case 'D': case 'd': return 13;
		case 'H': case 'h': return 11;
		case 'V': case 'v': return 7;
		case 'N': case 'n': return 15;
		case '-': return 0;
	}
	return 16;
}

// Match-mask planes of a byte string, for the per-slice masks of getEqVector
// (src/GraphAlignerBitvectorCommon.h:279-318) without 64 byte reads per slice:
// for every 64-position block b, planes[kPlaneWords * b + k] has bit l set when
// the letter at 64 b + l admits base k (A, C, G, T; readBaseSet) for k < 4, and
// when readBaseSet rejects it for k = 4. One block past the string's last is
// kept (zero) for the shifted read of a slice starting in the last block.
constexpr uint32_t kPlaneWords = 5;
GA_HHD uint64_t planeBlockCount(uint64_t bytes) { return bytes / 64 + 2; }
GA_HHD void planeBlock(const char* s, uint64_t n, uint64_t block, uint64_t* out)
{
	uint64_t m[kPlaneWords] = {0, 0, 0, 0, 0};
	for (uint32_t l = 0; l < 64; l++)
	{
		const uint64_t x = block * 64 + l;
		if (x >= n) break;
		const uint32_t set = readBaseSet((unsigned char)s[x]);
		const uint64_t bit = (uint64_t)1 << l;
		if (set == 16) m[4] |= bit;
		if (set & 1) m[0] |= bit;
		if (set & 2) m[1] |= bit;
		if (set & 4) m[2] |= bit;
		if (set & 8) m[3] |= bit;
	}
	for (uint32_t k = 0; k < kPlaneWords; k++) out[k] = m[k];
}
// Masks of rows x0 .. x0 + 63 (rows at or past x0 + rows match nothing); true
// when one of the rows holds a letter readBaseSet rejects.
GA_HD bool eqFromPlanes(const uint64_t* planes, uint64_t x0, uint32_t rows, Eq4& eq)
{
	const uint64_t* a = planes + (x0 >> 6) * kPlaneWords;
	const uint64_t* b = a + kPlaneWords;
	const uint32_t sh = (uint32_t)(x0 & 63);
	const uint64_t keep = rows >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << rows) - 1);
	uint64_t v[kPlaneWords];
	for (uint32_t k = 0; k < kPlaneWords; k++) v[k] = (sh == 0 ? a[k] : ((a[k] >> sh) | (b[k] << (64 - sh)))) & keep;
	for (uint32_t k = 0; k < 4; k++) eq.m[k] = v[k];
	return v[4] != 0;
}

// ---------------------------------------------------------------------------
// Lane-parallel helpers. Each takes uniform inputs and returns a uniform result.
// ---------------------------------------------------------------------------
#if defined(GA_THREAD_PER_EXTENSION)

// Rows 0..63 where column a scores below column b (lt) and above it (gt), as a
// value (outputs through references made the inlined code keep a stack slot):
// the result of differenceMasks (src/WordSlice.h:546-664), computed here as a sweep
// over the rows where the difference d(r) = a(r) - b(r) changes. d(-1) is the
// difference of the scores before row 0; at row r it moves by a's vertical
// delta minus b's, one of -2..2, so its sign is constant between two such rows.
// The sweep stops early once the rises (falls) still ahead cannot bring a
// negative (positive) difference back to zero.
struct RowCompare
{
	uint64_t lt, gt;
};
GA_HD RowCompare compareColumns(const Slice& a, const Slice& b)
{
	uint64_t lt, gt;
	// rows where d rises (a's +1 against b's 0/-1, or b's -1 against a's 0), where
	// it falls, and where it moves by two (+1 against -1)
	const uint64_t rise = (a.VP & ~b.VP) | (b.VN & ~a.VN);
	const uint64_t fall = (a.VN & ~b.VN) | (b.VP & ~a.VP);
	const uint64_t two = (a.VP & b.VN) | (a.VN & b.VP);
	int32_t d = before(a) - before(b);
	uint64_t events = rise | fall; // rows not swept yet where d changes
	lt = 0;
	gt = 0;
	uint32_t start = 0; // first row whose difference is d
	while (true)
	{
		// the falls (rises) left move d by at most two each
		if (d > 0 && d > 2 * popc(fall & events))
		{
			gt |= ONES << start;
			return RowCompare{lt, gt};
		}
		if (d < 0 && -d > 2 * popc(rise & events))
		{
			lt |= ONES << start;
			return RowCompare{lt, gt};
		}
		const uint32_t r = events ? (uint32_t)__builtin_ctzll(events) : 64;
		const uint64_t span = (r >= 64 ? ONES : (((uint64_t)1 << r) - 1)) & (ONES << start);
		if (d > 0) gt |= span;
		else if (d < 0) lt |= span;
		if (r >= 64) return RowCompare{lt, gt};
		const uint64_t bit = (uint64_t)1 << r;
		const int32_t step = (two & bit) ? 2 : 1;
		d += (rise & bit) ? step : -step;
		events &= events - 1;
		start = r;
	}
}

// Cellwise minimum of two columns (WordSlice::mergeWith, src/WordSlice.h:171-175),
// rebuilt from the comparison masks: where the same column is the smaller (ties
// to a) at rows r - 1 and r, row r keeps that column's vertical delta; where the
// smaller column switches between two rows that differ by a strict inequality
// on both sides, the minimum does not move (the deltas are confined to -1..1);
// where it switches across a tie, row r takes the delta of the column that is
// (also) smallest at both rows, b.
GA_HD Slice lpMerge(const Slice& a, const Slice& b)
{
	const RowCompare rc = compareColumns(a, b);
	const uint64_t lt = rc.lt, gt = rc.gt;
	const int32_t ba = before(a), bb = before(b);
	const uint64_t eq = ~(lt | gt);
	const uint64_t gtPrev = (gt << 1) | (uint64_t)(bb < ba);
	const uint64_t ltPrev = (lt << 1) | (uint64_t)(ba < bb);
	const uint64_t useB = (gt & ~ltPrev) | (gtPrev & eq);
	const uint64_t flat = (gt & ltPrev) | (gtPrev & lt);
	const uint64_t useA = ~(useB | flat);
	Slice r;
	r.VP = (useA & a.VP) | (useB & b.VP);
	r.VN = (useA & a.VN) | (useB & b.VN);
	r.end = imin(a.end, b.end);
	return r;
}
GA_HD uint64_t lpLessMask(const Slice& a, const Slice& b)
{
	return compareColumns(a, b).lt;
}
// minimum over the rows of mask, walking from one selected row to the next
GA_HD int32_t lpMinOver(const Slice& s, uint64_t mask)
{
	int32_t b = before(s);
	int32_t r = SMAX;
	while (mask != 0)
	{
		uint64_t cur = mask ^ (mask - 1);
		r = imin(r, b + popc(s.VP & cur) - popc(s.VN & cur));
		mask &= ~cur;
	}
	return r;
}
GA_HD int32_t lpMaxX(const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	int32_t b = before(s);
	int32_t r = SMIN;
	while (cand != 0)
	{
		uint64_t cur = cand ^ (cand - 1);
		int32_t cells = popc(cur);
		r = imax(r, (seqOffset + cells) * 100 - (b + popc(s.VP & cur) - popc(s.VN & cur)) * cost);
		cand &= ~cur;
	}
	return r;
}
struct LaneAcc
{
	int32_t v;
};
GA_HD void accInit(LaneAcc& a) { a.v = SMIN; }
GA_HD void accAdd(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand) { a.v = imax(a.v, lpMaxX(s, seqOffset, cost, cand)); }
GA_HD int32_t accReduce(const LaneAcc& a) { return a.v; }
#if defined(GA_THREAD_HOST_TEST)
GA_HD void accAddAlone(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand) { accAdd(a, s, seqOffset, cost, cand); }
#else
// lpMaxX for a wave that runs its extension alone (a one-thread workgroup: the
// other 63 lanes never execute anything else). The block switches all 64 lanes
// on, lane r scores row r of the column, X(r) = (seqOffset + r + 1) * 100 -
// cost * value(r) with value(r) = before + popc(VP & rows 0..r) - popc(VN & rows
// 0..r), or INT32_MIN when r is not a candidate; a DPP maximum over rows of 16
// lanes, then across rows (row_bcast), leaves the maximum in lane 63; the lanes
// are switched back off before the block ends. The scalar loop costs about 16
// scalar instructions per candidate; this costs a fixed 36 vector ones that the
// otherwise idle vector unit issues. int32 arithmetic wraps as in lpMaxX.
GA_HD int32_t lpMaxXAlone(const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	const uint32_t vplo = (uint32_t)s.VP, vphi = (uint32_t)(s.VP >> 32);
	const uint32_t vnlo = (uint32_t)s.VN, vnhi = (uint32_t)(s.VN >> 32);
	const uint32_t clo = (uint32_t)cand, chi = (uint32_t)(cand >> 32);
	const int32_t bef = before(s);
	const int32_t base = (seqOffset + 1) * 100;
	int32_t res;
	uint64_t saved;
	uint32_t ln, mlo, mhi, t0, t1, t2;
	asm volatile(
		"s_mov_b64 %[sv], exec\n\t"
		"s_mov_b64 exec, -1\n\t"
		// lane id, and the mask of rows 0..lane as two 32-bit halves
		"v_mbcnt_lo_u32_b32 %[ln], -1, 0\n\t"
		"v_mbcnt_hi_u32_b32 %[ln], -1, %[ln]\n\t"
		"v_lshlrev_b32_e64 %[t0], %[ln], 2\n\t"
		"v_add_u32_e32 %[t0], -1, %[t0]\n\t"
		"v_cmp_gt_u32_e32 vcc, 32, %[ln]\n\t"
		"v_cndmask_b32_e32 %[mlo], -1, %[t0], vcc\n\t"
		"v_cndmask_b32_e64 %[mhi], %[t0], 0, vcc\n\t"
		// value(lane)
		"v_and_b32_e32 %[t0], %[vplo], %[mlo]\n\t"
		"v_and_b32_e32 %[t1], %[vphi], %[mhi]\n\t"
		"v_bcnt_u32_b32 %[t0], %[t0], 0\n\t"
		"v_bcnt_u32_b32 %[t0], %[t1], %[t0]\n\t"
		"v_and_b32_e32 %[t1], %[vnlo], %[mlo]\n\t"
		"v_and_b32_e32 %[t2], %[vnhi], %[mhi]\n\t"
		"v_bcnt_u32_b32 %[t1], %[t1], 0\n\t"
		"v_bcnt_u32_b32 %[t1], %[t2], %[t1]\n\t"
		"v_sub_u32_e32 %[t0], %[t0], %[t1]\n\t"
		"v_add_u32_e32 %[t0], %[bef], %[t0]\n\t"
		// X(lane)
		"v_mul_lo_u32 %[t0], %[t0], %[cost]\n\t"
		"v_mul_u32_u24_e32 %[t1], 100, %[ln]\n\t"
		"v_add_u32_e32 %[t1], %[base], %[t1]\n\t"
		"v_sub_u32_e32 %[t0], %[t1], %[t0]\n\t"
		// INT32_MIN unless the lane's row is a candidate (vcc still holds lane < 32)
		"v_lshrrev_b32_e64 %[t1], %[ln], %[clo]\n\t"
		"v_lshrrev_b32_e64 %[t2], %[ln], %[chi]\n\t"
		"v_cndmask_b32_e32 %[t1], %[t2], %[t1], vcc\n\t"
		"v_and_b32_e32 %[t1], 1, %[t1]\n\t"
		"v_cmp_eq_u32_e32 vcc, 1, %[t1]\n\t"
		"v_bfrev_b32_e32 %[t2], 1\n\t"
		"v_cndmask_b32_e32 %[t0], %[t2], %[t0], vcc\n\t"
		// maximum over the wave into lane 63 (two wait states before each DPP read)
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_max_i32_dpp %[t0], %[t0], %[t0] row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
		"s_nop 1\n\t"
		"v_readlane_b32 %[res], %[t0], 63\n\t"
		"s_mov_b64 exec, %[sv]"
		: [res] "=s"(res), [sv] "=&s"(saved), [ln] "=&v"(ln), [mlo] "=&v"(mlo), [mhi] "=&v"(mhi), [t0] "=&v"(t0), [t1] "=&v"(t1),
		  [t2] "=&v"(t2)
		: [vplo] "s"(vplo), [vphi] "s"(vphi), [vnlo] "s"(vnlo), [vnhi] "s"(vnhi), [clo] "s"(clo), [chi] "s"(chi), [bef] "s"(bef),
		  [cost] "s"(cost), [base] "s"(base)
		: "vcc");
	return res;
}
GA_HD void accAddAlone(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand) { a.v = imax(a.v, lpMaxXAlone(s, seqOffset, cost, cand)); }
#endif
GA_HD int32_t lpFind(const uint32_t* arr, uint32_t n, uint32_t key)
{
	for (uint32_t i = 0; i < n; i++)
		if (arr[i] == key) return (int32_t)i;
	return -1;
}
GA_HD bool lpEq(const char* seq, uint32_t len, uint32_t j, Eq4& eq)
{
	eq.m[0] = eq.m[1] = eq.m[2] = eq.m[3] = 0;
	bool invalid = false;
	for (uint32_t l = 0; l < 64 && j + l < len; l++)
	{
		uint32_t set = readBaseSet((unsigned char)seq[j + l]);
		if (set == 16) invalid = true;
		uint64_t bit = (uint64_t)1 << l;
		if (set & 1) eq.m[0] |= bit;
		if (set & 2) eq.m[1] |= bit;
		if (set & 4) eq.m[2] |= bit;
		if (set & 8) eq.m[3] |= bit;
	}
	return invalid;
}
GA_HD uint32_t lpArgMin(const uint64_t* key, const uint32_t* seq, uint32_t n)
{
	uint32_t bi = 0;
	uint64_t bk = key[0];
	uint32_t bs = seq[0];
	for (uint32_t i = 1; i < n; i++)
	{
		uint64_t k = key[i];
		uint32_t s = seq[i];
		if (k < bk || (k == bk && s < bs))
		{
			bk = k;
			bs = s;
			bi = i;
		}
	}
	return bi;
}
GA_HD uint64_t mixHash(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)); }
GA_HD uint64_t itemHash(const Item& it)
{
	uint64_t h = 0x12345;
	h = mixHash(h, it.node);
	h = mixHash(h, it.sVP);
	h = mixHash(h, it.sVN);
	h = mixHash(h, (uint32_t)it.sEnd);
	h = mixHash(h, it.eVP);
	h = mixHash(h, it.eVN);
	h = mixHash(h, (uint32_t)it.eEnd);
	h = mixHash(h, it.HP);
	h = mixHash(h, it.HN);
	h = mixHash(h, (uint32_t)it.minScore);
	return h;
}
GA_HD uint64_t lpDigest(const Item* items, uint32_t start, uint32_t count)
{
	uint64_t h = 0;
	for (uint32_t i = 0; i < count; i++) h += itemHash(items[start + i]);
	return h;
}
// a node's column slices, in the thread's workspace or in LDS; with each column
// c >= 1 also the horizontal deltas (Ph, Mh) of the step from column c - 1 and
// the column's match mask
struct ColumnStore
{
	uint64_t* vp;
	uint64_t* vn;
	uint64_t* ph;
	uint64_t* mh;
	uint64_t* eq;
	int32_t* end;
	bool lds;
};
#if !defined(GA_THREAD_HOST_TEST)
// An LDS column store is accessed through LDS-typed pointers. Flat accesses to
// it would count on the vector-memory counter, so that every column read of the
// backtrace walk would also wait for all trace stores still in flight.
typedef __attribute__((address_space(3))) uint64_t LdsU64;
typedef __attribute__((address_space(3))) int32_t LdsI32;
GA_HD void colSet(ColumnStore& c, int col, const Slice& s)
{
	if (c.lds)
	{
		((LdsU64*)c.vp)[col] = s.VP;
		((LdsU64*)c.vn)[col] = s.VN;
		((LdsI32*)c.end)[col] = s.end;
		return;
	}
	c.vp[col] = s.VP;
	c.vn[col] = s.VN;
	c.end[col] = s.end;
}
GA_HD Slice colGet(const ColumnStore& c, int col)
{
	if (c.lds) return mkslice(((LdsU64*)c.vp)[col], ((LdsU64*)c.vn)[col], ((LdsI32*)c.end)[col]);
	return mkslice(c.vp[col], c.vn[col], c.end[col]);
}
GA_HD void colSetStep(ColumnStore& c, int col, uint64_t ph, uint64_t mh, uint64_t eq)
{
	if (c.lds)
	{
		((LdsU64*)c.ph)[col] = ph;
		((LdsU64*)c.mh)[col] = mh;
		((LdsU64*)c.eq)[col] = eq;
		return;
	}
	c.ph[col] = ph;
	c.mh[col] = mh;
	c.eq[col] = eq;
}
GA_HD ColStep colGetStep(const ColumnStore& c, int col)
{
	if (c.lds) return mkstep(((LdsU64*)c.ph)[col], ((LdsU64*)c.mh)[col], ((LdsU64*)c.eq)[col]);
	return mkstep(c.ph[col], c.mh[col], c.eq[col]);
}
#else
GA_HD void colSet(ColumnStore& c, int col, const Slice& s)
{
	c.vp[col] = s.VP;
	c.vn[col] = s.VN;
	c.end[col] = s.end;
}
GA_HD Slice colGet(const ColumnStore& c, int col) { return mkslice(c.vp[col], c.vn[col], c.end[col]); }
GA_HD void colSetStep(ColumnStore& c, int col, uint64_t ph, uint64_t mh, uint64_t eq)
{
	c.ph[col] = ph;
	c.mh[col] = mh;
	c.eq[col] = eq;
}
GA_HD ColStep colGetStep(const ColumnStore& c, int col) { return mkstep(c.ph[col], c.mh[col], c.eq[col]); }
#endif
GA_HD void bindColumns(ColumnStore& c, uint64_t* words, int32_t* end, bool lds)
{
	c.vp = words;
	c.vn = words + 64;
	c.ph = words + 128;
	c.mh = words + 192;
	c.eq = words + 256;
	c.end = end;
	c.lds = lds;
}

#elif !defined(GA_WAVE_EMULATION)

GA_HD int laneId() { return __lane_id(); }
GA_HD uint64_t prefixMask(int lane) { return lane == 63 ? ONES : (((uint64_t)1 << (lane + 1)) - 1); }
// value of row `lane` of slice s
GA_HD int32_t laneValue(const Slice& s, int32_t bef, uint64_t pm) { return bef + popc(s.VP & pm) - popc(s.VN & pm); }
GA_HD int32_t waveMaxI(int32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = imax(v, __shfl_xor(v, o));
	return v;
}
GA_HD int32_t waveMinI(int32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = imin(v, __shfl_xor(v, o));
	return v;
}
GA_HD uint64_t waveSumU64(uint64_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

// Cellwise minimum of two columns over rows -1..63, in slice form. Equals
// WordSlice::mergeTwoSlices (src/WordSlice.h:502-541) for valid slices.
GA_HD Slice lpMerge(const Slice& a, const Slice& b)
{
	int l = laneId();
	uint64_t pm = prefixMask(l);
	int32_t ba = before(a), bb = before(b);
	int32_t bm = imin(ba, bb);
	int32_t v = imin(laneValue(a, ba, pm), laneValue(b, bb, pm));
	int32_t up = __shfl_up(v, 1);
	if (l == 0) up = bm;
	int32_t d = v - up;
	Slice r;
	r.VP = __ballot(d == 1);
	r.VN = __ballot(d == -1);
	r.end = imin(a.end, b.end);
	return r;
}

// rows where a < b
GA_HD uint64_t lpLessMask(const Slice& a, const Slice& b)
{
	uint64_t pm = prefixMask(laneId());
	return __ballot(laneValue(a, before(a), pm) < laneValue(b, before(b), pm));
}

// minimum of the row values selected by mask (SMAX if empty)
GA_HD int32_t lpMinOver(const Slice& s, uint64_t mask)
{
	int l = laneId();
	int32_t v = ((mask >> l) & 1) ? laneValue(s, before(s), prefixMask(l)) : SMAX;
	return waveMinI(v);
}

// max over candidate rows r of (seqOffset + r + 1) * 100 - value(r) * cost
GA_HD int32_t lpMaxX(const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	int l = laneId();
	int32_t v = SMIN;
	if ((cand >> l) & 1) v = (seqOffset + l + 1) * 100 - laneValue(s, before(s), prefixMask(l)) * cost;
	return waveMaxI(v);
}

// per-lane running maximum of X-scores, reduced once per node
struct LaneAcc
{
	int32_t v;
};
GA_HD void accInit(LaneAcc& a) { a.v = SMIN; }
GA_HD void accAdd(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	int l = laneId();
	if ((cand >> l) & 1) a.v = imax(a.v, (seqOffset + l + 1) * 100 - laneValue(s, before(s), prefixMask(l)) * cost);
}
GA_HD int32_t accReduce(const LaneAcc& a) { return waveMaxI(a.v); }
GA_HD void accAddAlone(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand) { accAdd(a, s, seqOffset, cost, cand); }

// index of key in arr[0, n), or -1
GA_HD int32_t lpFind(const uint32_t* arr, uint32_t n, uint32_t key)
{
	int l = laneId();
	for (uint32_t base = 0; base < n; base += 64)
	{
		uint32_t i = base + l;
		bool hit = i < n && arr[i] == key;
		uint64_t m = __ballot(hit);
		if (m) return (int32_t)(base + __builtin_ctzll(m));
	}
	return -1;
}

// Eq vector of rows j..j+63 (src/GraphAlignerBitvectorCommon.h:279-318)
GA_HD bool lpEq(const char* seq, uint32_t len, uint32_t j, Eq4& eq)
{
	int l = laneId();
	uint32_t set = 0;
	if (j + l < len) set = readBaseSet((unsigned char)seq[j + l]);
	eq.m[0] = __ballot(set & 1);
	eq.m[1] = __ballot((set >> 1) & 1);
	eq.m[2] = __ballot((set >> 2) & 1);
	eq.m[3] = __ballot((set >> 3) & 1);
	return __ballot(set == 16) != 0;
}

// index of the smallest (key, seq) among n queue entries
GA_HD uint32_t lpArgMin(const uint64_t* key, const uint32_t* seq, uint32_t n)
{
	int l = laneId();
	uint64_t bk = ONES;
	uint32_t bs = NIL, bi = NIL;
	for (uint32_t i = l; i < n; i += 64)
	{
		uint64_t k = key[i];
		uint32_t s = seq[i];
		if (k < bk || (k == bk && s < bs))
		{
			bk = k;
			bs = s;
			bi = i;
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
	{
		uint64_t ok = __shfl_xor(bk, o);
		uint32_t os = __shfl_xor(bs, o);
		uint32_t oi = __shfl_xor(bi, o);
		if (ok < bk || (ok == bk && os < bs))
		{
			bk = ok;
			bs = os;
			bi = oi;
		}
	}
	return bi;
}

GA_HD uint64_t mixHash(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)); }
GA_HD uint64_t itemHash(const Item& it)
{
	uint64_t h = 0x12345;
	h = mixHash(h, it.node);
	h = mixHash(h, it.sVP);
	h = mixHash(h, it.sVN);
	h = mixHash(h, (uint32_t)it.sEnd);
	h = mixHash(h, it.eVP);
	h = mixHash(h, it.eVN);
	h = mixHash(h, (uint32_t)it.eEnd);
	h = mixHash(h, it.HP);
	h = mixHash(h, it.HN);
	h = mixHash(h, (uint32_t)it.minScore);
	return h;
}
// order-independent digest of a slice's band items (test output only)
GA_HD uint64_t lpDigest(const Item* items, uint32_t start, uint32_t count)
{
	uint64_t h = 0;
	for (uint32_t i = laneId(); i < count; i += 64) h += itemHash(items[start + i]);
	return waveSumU64(h);
}

// the node's column slices, column c held by lane c
struct ColumnStore
{
	uint64_t vp, vn, ph, mh, eq;
	int32_t end;
};
GA_HD void colSet(ColumnStore& c, int col, const Slice& s)
{
	if (laneId() == col)
	{
		c.vp = s.VP;
		c.vn = s.VN;
		c.end = s.end;
	}
}
GA_HD void colSetStep(ColumnStore& c, int col, uint64_t ph, uint64_t mh, uint64_t eq)
{
	if (laneId() == col)
	{
		c.ph = ph;
		c.mh = mh;
		c.eq = eq;
	}
}
GA_HD ColStep colGetStep(const ColumnStore& c, int col) { return mkstep(__shfl(c.ph, col), __shfl(c.mh, col), __shfl(c.eq, col)); }
GA_HD Slice colGet(const ColumnStore& c, int col)
{
	Slice s;
	s.VP = __shfl(c.vp, col);
	s.VN = __shfl(c.vn, col);
	s.end = __shfl(c.end, col);
	return s;
}
GA_HD void bindColumns(ColumnStore&, uint64_t*, int32_t*, bool) {}

#else // GA_WAVE_EMULATION: the same helpers as loops over 64 lanes

inline uint64_t prefixMask(int lane) { return lane == 63 ? ONES : (((uint64_t)1 << (lane + 1)) - 1); }
inline int32_t laneValue(const Slice& s, int32_t bef, uint64_t pm) { return bef + popc(s.VP & pm) - popc(s.VN & pm); }
inline Slice lpMerge(const Slice& a, const Slice& b)
{
	int32_t ba = before(a), bb = before(b);
	int32_t up = imin(ba, bb);
	Slice r;
	r.VP = 0;
	r.VN = 0;
	for (int l = 0; l < 64; l++)
	{
		uint64_t pm = prefixMask(l);
		int32_t v = imin(laneValue(a, ba, pm), laneValue(b, bb, pm));
		if (v - up == 1) r.VP |= (uint64_t)1 << l;
		if (v - up == -1) r.VN |= (uint64_t)1 << l;
		up = v;
	}
	r.end = imin(a.end, b.end);
	return r;
}
inline uint64_t lpLessMask(const Slice& a, const Slice& b)
{
	uint64_t m = 0;
	int32_t ba = before(a), bb = before(b);
	for (int l = 0; l < 64; l++)
		if (laneValue(a, ba, prefixMask(l)) < laneValue(b, bb, prefixMask(l))) m |= (uint64_t)1 << l;
	return m;
}
inline int32_t lpMinOver(const Slice& s, uint64_t mask)
{
	int32_t r = SMAX;
	for (int l = 0; l < 64; l++)
		if ((mask >> l) & 1) r = imin(r, laneValue(s, before(s), prefixMask(l)));
	return r;
}
inline int32_t lpMaxX(const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	int32_t r = SMIN;
	for (int l = 0; l < 64; l++)
		if ((cand >> l) & 1) r = imax(r, (seqOffset + l + 1) * 100 - laneValue(s, before(s), prefixMask(l)) * cost);
	return r;
}
struct LaneAcc
{
	int32_t v[64];
};
inline void accInit(LaneAcc& a)
{
	for (int l = 0; l < 64; l++) a.v[l] = SMIN;
}
inline void accAdd(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand)
{
	for (int l = 0; l < 64; l++)
		if ((cand >> l) & 1) a.v[l] = imax(a.v[l], (seqOffset + l + 1) * 100 - laneValue(s, before(s), prefixMask(l)) * cost);
}
inline int32_t accReduce(const LaneAcc& a)
{
	int32_t r = SMIN;
	for (int l = 0; l < 64; l++) r = imax(r, a.v[l]);
	return r;
}
inline void accAddAlone(LaneAcc& a, const Slice& s, int32_t seqOffset, int32_t cost, uint64_t cand) { accAdd(a, s, seqOffset, cost, cand); }
inline int32_t lpFind(const uint32_t* arr, uint32_t n, uint32_t key)
{
	for (uint32_t i = 0; i < n; i++)
		if (arr[i] == key) return (int32_t)i;
	return -1;
}
inline bool lpEq(const char* seq, uint32_t len, uint32_t j, Eq4& eq)
{
	eq.m[0] = eq.m[1] = eq.m[2] = eq.m[3] = 0;
	bool invalid = false;
	for (int l = 0; l < 64; l++)
	{
		if (j + l >= len) continue;
		uint32_t set = readBaseSet((unsigned char)seq[j + l]);
		if (set == 16) invalid = true;
		for (int b = 0; b < 4; b++)
			if ((set >> b) & 1) eq.m[b] |= (uint64_t)1 << l;
	}
	return invalid;
}
inline uint32_t lpArgMin(const uint64_t* key, const uint32_t* seq, uint32_t n)
{
	uint32_t bi = 0;
	for (uint32_t i = 1; i < n; i++)
		if (key[i] < key[bi] || (key[i] == key[bi] && seq[i] < seq[bi])) bi = i;
	return bi;
}
GA_HD uint64_t mixHash(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)); }
GA_HD uint64_t itemHash(const Item& it)
{
	uint64_t h = 0x12345;
	h = mixHash(h, it.node);
	h = mixHash(h, it.sVP);
	h = mixHash(h, it.sVN);
	h = mixHash(h, (uint32_t)it.sEnd);
	h = mixHash(h, it.eVP);
	h = mixHash(h, it.eVN);
	h = mixHash(h, (uint32_t)it.eEnd);
	h = mixHash(h, it.HP);
	h = mixHash(h, it.HN);
	h = mixHash(h, (uint32_t)it.minScore);
	return h;
}
inline uint64_t lpDigest(const Item* items, uint32_t start, uint32_t count)
{
	uint64_t h = 0;
	for (uint32_t i = 0; i < count; i++) h += itemHash(items[start + i]);
	return h;
}
struct ColumnStore
{
	Slice c[64];
	ColStep st[64];
};
inline void colSet(ColumnStore& c, int col, const Slice& s) { c.c[col] = s; }
inline Slice colGet(const ColumnStore& c, int col) { return c.c[col]; }
inline void colSetStep(ColumnStore& c, int col, uint64_t ph, uint64_t mh, uint64_t eq) { c.st[col] = mkstep(ph, mh, eq); }
inline ColStep colGetStep(const ColumnStore& c, int col) { return c.st[col]; }
inline void bindColumns(ColumnStore&, uint64_t*, int32_t*, bool) {}

#endif

// ---------------------------------------------------------------------------
// One extension
// ---------------------------------------------------------------------------
struct NodeResult
{
	int32_t minScore;
	int32_t minScoreNodeOffset;
	int32_t maxX;
	int32_t cells;
};

// Band queue entry key: component in the high word, score (order-preserving) low.
GA_HD uint64_t queueKey(uint32_t comp, int32_t score) { return ((uint64_t)comp << 32) | (uint32_t)((uint32_t)score ^ 0x80000000u); }

GA_HD uint64_t packTrace(uint32_t node, int32_t seqPos, uint32_t offset, uint32_t nodeSwitch)
{
	uint32_t hi = ((uint32_t)(seqPos + 1) << 7) | (offset << 1) | nodeSwitch;
	return ((uint64_t)hi << 32) | node;
}

struct Pos
{
	uint32_t node;
	uint32_t off;
	int32_t seq;
};

GA_HD bool samePos(const Pos& a, const Pos& b) { return a.node == b.node && a.off == b.off && a.seq == b.seq; }
GA_HD Pos mkpos(uint32_t node, uint32_t off, int32_t seq)
{
	Pos p;
	p.node = node;
	p.off = off;
	p.seq = seq;
	return p;
}

struct Extension
{
	// inputs
	DevGraph g;
	ExtParams p;
	const char* seq;
	uint32_t len;
	uint32_t startNode;
	uint32_t startOffset;
	// per-slice bandwidth schedule (getXdropSlices picks a bandwidth per slice,
	// src/GraphAlignerBitvectorBanded.h:611): slice s uses bands[min(s, bandLen - 1)];
	// bandLen 0 uses p.bandwidth
	const int32_t* bands = nullptr;
	uint32_t bandLen = 0;
	// the wave runs this extension alone (a one-thread workgroup): its other
	// lanes may be switched on for whole-wave steps (lpMaxXAlone)
	bool aloneInWave = false;
	// match-mask planes of the buffer seq points into (planeBlock) and seq's
	// offset there; without planes the masks are read from seq byte by byte
	const uint64_t* planes;
	uint64_t seqBase;
	// workspace
	ArenaCaps caps;
	SliceHdr* hdr;
	Item* items;
	uint32_t* itemNode;
	uint32_t* slotNode;
	int32_t* slotItem;
	uint32_t* slotHead;
	uint32_t* slotTail;
	uint32_t* slotActive;
	uint64_t* exVP;
	uint64_t* exVN;
	int32_t* exEnd;
	uint32_t* exNext;
	uint64_t* pendKey;
	uint32_t* pendSeq;
	uint32_t* pendSlot;
	uint64_t* colWords; // kColumnWords: VP, VN, Ph, Mh, Eq of 64 columns, then 64 scores
	int32_t* colEnd;
	bool colsInLds; // colWords/colEnd point into LDS
	// node -> item index of every kept slice: one open-addressing table per slice
	// (entries node << 32 | item, ~0 empty), sized to a power of two at least
	// twice the slice's item count; its first entry is the header's pad field
	uint64_t* sliceTab;
	uint32_t sliceTabCap, sliceTabUsed;
	// node -> slot of the slice being computed: entries stamp << 32 | node, with
	// the slot beside them; entries of earlier slices carry older stamps and
	// count as empty, so the table is never cleared
	uint64_t* slotTabKey;
	uint32_t* slotTabVal;
	uint32_t slotTabMask;
	uint32_t slotStamp;
	// forward pass: the items of the slice being computed (index = item - the
	// slice's first item) and of the previous kept slice, with a node -> index
	// table of the latter (entries node << 32 | index, ~0 empty); the kept
	// slices' items are copied to `items` when their slice ends
	Item* curItems;
	Item* prevItems;
	uint64_t* prevTab;
	uint32_t prevTabMask;
	// outputs
	uint64_t* trace;
	uint32_t traceCap;
	SliceSummary* summaries; // optional
	// state
	uint32_t nItems, kept, nSlots, nEx, nPend, seqCounter;
	uint32_t ntrace;
	int32_t status;
	uint32_t overflowWhat;
	uint64_t wordSteps, nodeCalcs;
	uint32_t maxBand;
	Pos back;

	GA_HD uint32_t nodeLen(uint32_t n) const { return g.rec[n].len; }
	GA_HD uint32_t nodeBaseSet(uint32_t n, uint32_t pos) const
	{
		if (n < g.firstAmbiguous)
		{
			uint32_t code = (uint32_t)(((pos >> 5) ? g.rec[n].seq1 : g.rec[n].seq0) >> ((pos & 31) * 2)) & 3;
			return 1u << code;
		}
		const uint64_t* m = g.amb4 + 4 * (uint64_t)(n - g.firstAmbiguous);
		return (uint32_t)((m[0] >> pos) & 1) | (uint32_t)(((m[1] >> pos) & 1) << 1) | (uint32_t)(((m[2] >> pos) & 1) << 2) | (uint32_t)(((m[3] >> pos) & 1) << 3);
	}
	GA_HD uint64_t eqAt(const Eq4& eq, uint32_t n, uint32_t pos) const
	{
		if (n < g.firstAmbiguous)
		{
			uint32_t code = (uint32_t)(((pos >> 5) ? g.rec[n].seq1 : g.rec[n].seq0) >> ((pos & 31) * 2)) & 3;
			return eqSelect(eq, code);
		}
		uint32_t set = nodeBaseSet(n, pos);
		uint64_t r = 0;
		if (set & 1) r |= eq.m[0];
		if (set & 2) r |= eq.m[1];
		if (set & 4) r |= eq.m[2];
		if (set & 8) r |= eq.m[3];
		return r;
	}
	// A node's bases held in registers for its column loop: two 2-bit chunks, or
	// the four base masks of an IUPAC node.
	struct NodeSeq
	{
		uint64_t w[4];
		bool amb;
	};
	GA_HD NodeSeq loadNodeSeq(uint32_t n) const
	{
		NodeSeq s;
		s.amb = n >= g.firstAmbiguous;
		if (!s.amb)
		{
			s.w[0] = g.rec[n].seq0;
			s.w[1] = g.rec[n].seq1;
			s.w[2] = s.w[3] = 0;
		}
		else
		{
			const uint64_t* m = g.amb4 + 4 * (uint64_t)(n - g.firstAmbiguous);
			s.w[0] = m[0];
			s.w[1] = m[1];
			s.w[2] = m[2];
			s.w[3] = m[3];
		}
		return s;
	}
	// eqAt from registers
	GA_HD uint64_t eqOf(const Eq4& eq, const NodeSeq& s, uint32_t pos) const
	{
		if (!s.amb)
		{
			uint64_t w = pos < 32 ? s.w[0] : s.w[1];
			return eqSelect(eq, (uint32_t)(w >> ((pos & 31) * 2)) & 3);
		}
		uint64_t r = 0;
		if ((s.w[0] >> pos) & 1) r |= eq.m[0];
		if ((s.w[1] >> pos) & 1) r |= eq.m[1];
		if ((s.w[2] >> pos) & 1) r |= eq.m[2];
		if ((s.w[3] >> pos) & 1) r |= eq.m[3];
		return r;
	}
	// characterMatch(read char, graph char) (src/GraphAlignerCommon.h:176-198)
	// match masks of rows j .. j + 63 (getEqVector); true on a letter readBaseSet rejects
	GA_HD bool sliceEq(uint32_t j, Eq4& eq) const
	{
		if (planes) return eqFromPlanes(planes, seqBase + j, len - j, eq);
		return lpEq(seq, len, j, eq);
	}
	GA_HD bool charMatch(int32_t seqPos, uint32_t node, uint32_t off) const
	{
		uint32_t rs = readBaseSet((unsigned char)seq[seqPos]);
		return (rs & nodeBaseSet(node, off) & 15) != 0;
	}
	GA_HD void fail(int32_t st, uint32_t what = 0)
	{
		if (status == EXT_OK)
		{
			status = st;
			overflowWhat = what;
		}
	}

	// --- band queue (ComponentPriorityQueue with insertion-order ties) ---
	GA_HD static uint32_t nodeHash(uint32_t node) { return (node * 2654435761u) ^ (node >> 15); }
	GA_HD int32_t findSlot(uint32_t node) const
	{
		const uint64_t want = ((uint64_t)slotStamp << 32) | node;
		for (uint32_t i = nodeHash(node) & slotTabMask, probes = 0; probes <= slotTabMask; i = (i + 1) & slotTabMask, probes++)
		{
			uint64_t k = slotTabKey[i];
			if ((uint32_t)(k >> 32) != slotStamp) return -1;
			if (k == want) return (int32_t)slotTabVal[i];
		}
		return -1;
	}
	GA_HD void addSlotIndex(uint32_t node, uint32_t slot)
	{
		const uint64_t want = ((uint64_t)slotStamp << 32) | node;
		for (uint32_t i = nodeHash(node) & slotTabMask, probes = 0; probes <= slotTabMask; i = (i + 1) & slotTabMask, probes++)
		{
			if ((uint32_t)(slotTabKey[i] >> 32) != slotStamp)
			{
				slotTabKey[i] = want;
				slotTabVal[i] = slot;
				return;
			}
		}
		fail(EXT_OVERFLOW, 10);
	}

	GA_HD void queueInsert(uint32_t comp, int32_t score, uint32_t target, const Slice& s, bool skipFirst)
	{
		int32_t slot = findSlot(target);
		if (slot < 0)
		{
			if (nSlots >= caps.maxSlots) { fail(EXT_OVERFLOW, 2); return; }
			slot = nSlots++;
			addSlotIndex(target, (uint32_t)slot);
			slotNode[slot] = target;
			slotItem[slot] = -1;
			slotHead[slot] = NIL;
			slotTail[slot] = NIL;
			slotActive[slot] = 0;
		}
		if (nEx >= caps.maxExtras) { fail(EXT_OVERFLOW, 3); return; }
		if (!slotActive[slot])
		{
			if (nPend >= caps.maxPending) { fail(EXT_OVERFLOW, 4); return; }
			pendKey[nPend] = queueKey(comp, score);
			pendSeq[nPend] = seqCounter++;
			pendSlot[nPend] = (uint32_t)slot;
			nPend++;
			slotActive[slot] = 1;
		}
		uint32_t e = nEx++;
		exVP[e] = s.VP;
		exVN[e] = s.VN;
		exEnd[e] = s.end;
		exNext[e] = (skipFirst ? SKIPFIRST : 0) | LISTEND;
		uint32_t tail = slotTail[slot];
		if (tail == NIL) slotHead[slot] = e;
		else exNext[tail] = (exNext[tail] & SKIPFIRST) | e;
		slotTail[slot] = e;
	}

	// --- one node's columns for one slice (calculateNodeInner) ---
	// Forward mode (Recalc = false) with early leave; recalc mode stores every
	// column in `cols` and leaves the X-scores out.
	// xFloor: the slice's best X-score so far. The node's maxX is exact whenever
	// it exceeds xFloor (the only case where the slice uses it); columns whose
	// X-scores cannot exceed max(xFloor, column 0's) are not scanned.
	template <bool Recalc>
	GA_HD NodeResult calcNode(uint32_t node, Item& it, bool exists, const Eq4& eq, const Item& prev, bool prevExists, uint32_t firstExtra, const Slice& recalcIncoming, int32_t seqOffset, bool prevHasInNeighbor, ColumnStore* cols, int32_t xFloor = SMIN, uint32_t recalcCols = NIL)
	{
		NodeResult res;
		res.minScore = SMAX;
		res.minScoreNodeOffset = -1;
		res.maxX = SMIN;
		res.cells = 0;
		uint32_t L = nodeLen(node);
		const NodeSeq nseq = loadNodeSeq(node);
		uint64_t Eq0 = eqOf(eq, nseq, 0);
		Slice ws = mkslice(0, 0, 0);
		bool hasWs = false, hasSkipless = false;
		if (Recalc)
		{
			res.cells = 1;
			ws = recalcIncoming;
			hasWs = true;
		}
		else
		{
			for (uint32_t e = firstExtra; e < caps.maxExtras; e = exNext[e] & ~SKIPFIRST)
			{
				res.cells++;
				Slice inc = mkslice(exVP[e], exVN[e], exEnd[e]);
				if (exNext[e] & SKIPFIRST)
				{
					ws = hasWs ? lpMerge(ws, inc) : inc;
					hasWs = true;
					continue;
				}
				hasSkipless = true;
				int32_t compareHin = prevExists ? prev.sEnd : SMAX;
				int32_t ib = before(inc);
				uint64_t hinP = 0, hinN = 0;
				if (compareHin < ib) hinN = 1;
				else if (compareHin > ib) hinP = 1;
				uint64_t Ph, Mh;
				Slice stepped = nextSlice(Eq0, inc, hinP, hinN, Ph, Mh);
				if (!Recalc) wordSteps++;
				if (!prevExists || before(stepped) < prev.sEnd)
				{
					stepped.VP &= ONES ^ 1;
					stepped.VN |= 1;
				}
				ws = hasWs ? lpMerge(ws, stepped) : stepped;
				hasWs = true;
			}
		}
		res.minScore = ws.end;
		res.minScoreNodeOffset = 0;
		if (!Recalc) res.maxX = lpMaxX(ws, seqOffset, p.errorCost, maxXCandidates(ws, 64, ONES));

		if (exists)
		{
			Slice start = mkslice(it.sVP, it.sVN, it.sEnd);
			if (hasSkipless && g.rec[node].inCount == 1 && prevHasInNeighbor)
			{
				if (ws.end > start.end)
				{
					if (!Recalc) return res;
				}
				else if (ws.end < start.end)
				{
				}
				else
				{
					uint64_t newBigger = (ws.VP & ~start.VP) | (start.VN & ~ws.VN);
					uint64_t oldBigger = (start.VP & ~ws.VP) | (ws.VN & ~start.VN);
					if (newBigger > oldBigger)
					{
					}
					else if (oldBigger > newBigger)
					{
						if (!Recalc) return res;
					}
					else if (newBigger == 0 && oldBigger == 0)
					{
						if (!Recalc) return res;
					}
					else
					{
						Slice test = lpMerge(ws, start);
						if (same(test, start))
						{
							if (!Recalc) return res;
						}
						ws = test;
					}
				}
			}
			else
			{
				Slice test = lpMerge(ws, start);
				// src/GraphAlignerBitvectorCommon.h:1282 compares test.VP with startSlice.VN
				if (test.end == start.end && test.VP == start.VP && test.VP == start.VN)
				{
					if (!Recalc) return res;
				}
				ws = test;
			}
		}

		if (prevExists && before(ws) > prev.sEnd) ws = lpMerge(ws, sourceSlice(prev.sEnd));

		uint64_t forceMask = 0;
		uint64_t fixedHP = prev.HP, fixedHN = prev.HN;
		if (prevExists)
		{
			int32_t scoreBefore = before(ws);
			int32_t cmp = prev.sEnd;
			if (scoreBefore > cmp)
			{
				fail(EXT_ASSERT);
				return res;
			}
			if (scoreBefore < cmp)
			{
				forceMask |= 1;
				for (uint32_t fo = 1; fo < 64; fo++)
				{
					int32_t ncmp = cmp + (int32_t)((fixedHP >> fo) & 1) - (int32_t)((fixedHN >> fo) & 1);
					uint64_t mask = (uint64_t)1 << fo;
					if (scoreBefore > ncmp)
					{
						fail(EXT_ASSERT);
						return res;
					}
					if (scoreBefore < ncmp)
					{
						fixedHP |= mask;
						fixedHN &= ~mask;
						forceMask |= mask;
					}
					if (scoreBefore == ncmp)
					{
						fixedHP &= ~mask;
						fixedHN &= ~mask;
					}
					scoreBefore++;
					cmp = ncmp;
					if (scoreBefore >= cmp) break;
				}
			}
		}
		else
		{
			forceMask = ONES;
		}
		it.sVP = ws.VP;
		it.sVN = ws.VN;
		it.sEnd = ws.end;
		if (Recalc) colSet(*cols, 0, ws);
		uint64_t forceEq = prevExists ? ONES : (ONES ^ 1);
		uint64_t HP = 0, HN = 0;
		ws = mkslice(uni64(ws.VP), uni64(ws.VN), uni32s(ws.end));
		forceMask = uni64(forceMask);
		fixedHP = uni64(fixedHP);
		fixedHN = uni64(fixedHN);
		forceEq = uni64(forceEq);
		L = uni32(L);
		// row 0 never matches in a node without a previous-slice item (forceEq):
		// applied to the four base masks once instead of to every column's mask
		Eq4 eqU;
		for (int b = 0; b < 4; b++) eqU.m[b] = uni64(eq.m[b]) & forceEq;
		NodeSeq nseqU = nseq;
		for (int b = 0; b < 4; b++) nseqU.w[b] = uni64(nseq.w[b]);
		LaneAcc acc;
		// Row r of a column scores X(r) = (seqOffset + r + 1) * 100 - cost * value(r),
		// so X(r) - X(63) sums -100 + cost * (VP_i - VN_i) over the rows i > r: at
		// most X(63) + max(cost - 100, 0) * popc(VP) for every row. X-scores are
		// int32 as in lpMaxX.
		const int32_t xTop = (seqOffset + 64) * 100;
		const int32_t xRise = p.errorCost > 100 ? p.errorCost - 100 : 0;
		const int32_t xSkip = imax(xFloor, res.maxX);
		if (!Recalc)
		{
			accInit(acc);
		}
		// a recalculation may stop after the columns the backtrace will read
		const uint32_t Lc = Recalc && recalcCols < L ? recalcCols : L;
		// column pos (>= 1) of the node from its match mask
		auto column = [&](uint32_t pos, uint64_t Eq) {
			uint64_t Ph, Mh;
			Slice nw = nextSlice(Eq, ws, (fixedHP >> pos) & 1, (fixedHN >> pos) & 1, Ph, Mh);
			// row 0 forced to a vertical -1 where forceMask says so, with masks (a
			// branch on the bit compiles into a round trip through a vector register)
			const uint64_t f = (forceMask >> pos) & 1;
			nw.VP &= ~f;
			nw.VN |= f;
			ws = nw;
			if (ws.end < res.minScore)
			{
				res.minScore = ws.end;
				res.minScoreNodeOffset = pos;
			}
			if (Recalc)
			{
				colSet(*cols, pos, ws);
				colSetStep(*cols, pos, Ph, Mh, Eq);
			}
			else if (xTop - p.errorCost * ws.end + xRise * popc(ws.VP) > xSkip)
			{
				if (aloneInWave) accAddAlone(acc, ws, seqOffset, p.errorCost, maxXCandidates(ws, 64, Mh));
				else accAdd(acc, ws, seqOffset, p.errorCost, maxXCandidates(ws, 64, Mh));
			}
			HP |= (Ph >> 63) << pos;
			HN |= (Mh >> 63) << pos;
		};
		if (!nseqU.amb)
		{
			// the 2-bit bases as a shift register (column pos's code in the low two
			// bits), one loop per 32-base chunk: no per-column choice of the chunk
			// and no per-column test for IUPAC nodes
			uint64_t w = nseqU.w[0] >> 2;
			const uint32_t mid = Lc < 32 ? Lc : 32;
			for (uint32_t pos = 1; pos < mid; pos++)
			{
				column(pos, eqSelect(eqU, (uint32_t)w & 3));
				w >>= 2;
			}
			w = nseqU.w[1];
			for (uint32_t pos = 32; pos < Lc; pos++)
			{
				column(pos, eqSelect(eqU, (uint32_t)w & 3));
				w >>= 2;
			}
		}
		else
		{
			for (uint32_t pos = 1; pos < Lc; pos++) column(pos, eqOf(eqU, nseqU, pos));
		}
		if (!Recalc)
		{
			wordSteps += L - 1;
			if (L > 1) res.maxX = imax(res.maxX, accReduce(acc));
		}
		res.cells = L;
		it.HP = HP;
		it.HN = HN;
		it.eVP = ws.VP;
		it.eVN = ws.VN;
		it.eEnd = ws.end;
		if (Lc == L && prevExists && before(ws) > prev.eEnd) fail(EXT_ASSERT);
		return res;
	}

	// WordSlice::changedMinScore (src/WordSlice.h:438-472)
	GA_HD int32_t changedMinScore(const Slice& s, const Slice& old) const
	{
		int32_t b = before(s), ob = before(old);
		uint64_t m = (s.VP & (s.VN - s.VP)) >> 1;
		m |= LASTBIT & (s.VN | ~(s.VN - s.VP)) & ~s.VP;
		if (s.end + popc(s.VN) >= old.end - popc(old.VP))
		{
			uint64_t smaller = lpLessMask(s, old);
			if (smaller != ONES)
			{
				m |= (~smaller) >> 1;
				m |= (~smaller) << 1;
				m |= 1;
				m |= LASTBIT;
				m &= smaller;
			}
		}
		int32_t r = b < ob ? b : SMAX;
		return imin(r, lpMinOver(s, m));
	}

	GA_HD Item loadItem(uint32_t i) const { return items[i]; }
	GA_HD void storeItem(uint32_t i, const Item& it) { items[i] = it; }
	GA_HD static uint32_t sliceTabSize(uint32_t count)
	{
		uint32_t t = 4;
		while (t < 2 * count) t <<= 1;
		return t;
	}
	GA_HD int32_t findInSlice(uint32_t sliceIdx, uint32_t node) const { return findInSlice(hdr[sliceIdx], node); }
	// with the slice's header at hand (the forward pass keeps the previous one)
	GA_HD int32_t findInSlice(const SliceHdr& h, uint32_t node) const
	{
		if (h.itemCount == 0) return -1;
		const uint32_t mask = sliceTabSize(h.itemCount) - 1;
		const uint64_t* t = sliceTab + h.pad;
		for (uint32_t i = nodeHash(node) & mask, probes = 0; probes <= mask; i = (i + 1) & mask, probes++)
		{
			uint64_t e = t[i];
			if (e == ~(uint64_t)0) return -1;
			if ((uint32_t)(e >> 32) == node) return (int32_t)(uint32_t)e;
		}
		return -1;
	}
	// index of node's item in prevItems, or -1
	GA_HD int32_t findPrev(uint32_t node) const
	{
		for (uint32_t i = nodeHash(node) & prevTabMask, probes = 0; probes <= prevTabMask; i = (i + 1) & prevTabMask, probes++)
		{
			uint64_t e = prevTab[i];
			if (e == ~(uint64_t)0) return -1;
			if ((uint32_t)(e >> 32) == node) return (int32_t)(uint32_t)e;
		}
		return -1;
	}
	// indexes prevItems[0, count) for findPrev
	GA_HD void indexPrev(uint32_t count)
	{
		for (uint32_t i = 0; i <= prevTabMask; i++) prevTab[i] = ~(uint64_t)0;
		for (uint32_t k = 0; k < count; k++)
		{
			const uint32_t node = prevItems[k].node;
			uint32_t i = nodeHash(node) & prevTabMask;
			while (prevTab[i] != ~(uint64_t)0) i = (i + 1) & prevTabMask;
			prevTab[i] = ((uint64_t)node << 32) | k;
		}
	}
	// indexes the items of kept slice sliceIdx for findInSlice
	GA_HD void buildSliceTable(uint32_t sliceIdx)
	{
		SliceHdr& h = hdr[sliceIdx];
		const uint32_t size = sliceTabSize(h.itemCount);
		if (sliceTabUsed + size > sliceTabCap)
		{
			fail(EXT_OVERFLOW, 9);
			return;
		}
		h.pad = sliceTabUsed;
		uint64_t* t = sliceTab + sliceTabUsed;
		sliceTabUsed += size;
		for (uint32_t i = 0; i < size; i++) t[i] = ~(uint64_t)0;
		const uint32_t mask = size - 1;
		for (uint32_t k = 0; k < h.itemCount; k++)
		{
			uint32_t item = h.itemStart + k;
			uint32_t node = itemNode[item];
			uint32_t i = nodeHash(node) & mask;
			while (t[i] != ~(uint64_t)0) i = (i + 1) & mask;
			t[i] = ((uint64_t)node << 32) | item;
		}
	}
	GA_HD static Item defaultPrev()
	{
		Item d;
		d.sVP = d.sVN = d.eVP = d.eVN = 0;
		d.HP = ONES;
		d.HN = 0;
		d.sEnd = d.eEnd = d.minScore = 0;
		d.node = NIL;
		return d;
	}

	// --- forward pass: getXdropSlices + calculateSlice ---
	GA_HD void forward()
	{
		uint32_t L0 = nodeLen(startNode);
		Item first;
		first.node = startNode;
		first.sVP = first.sVN = first.eVP = first.eVN = 0;
		first.sEnd = (int32_t)startOffset;
		first.eEnd = (int32_t)L0 - 1 - (int32_t)startOffset;
		first.HP = 0;
		first.HN = 0;
		for (uint32_t i = 1; i <= startOffset; i++) first.HN |= (uint64_t)1 << i;
		for (uint32_t i = startOffset + 1; i < L0; i++) first.HP |= (uint64_t)1 << i;
		first.minScore = 0;
		items[0] = first;
		itemNode[0] = startNode;
		prevItems[0] = first;
		indexPrev(1);
		SliceHdr h0;
		h0.minScore = 0;
		h0.minScoreNode = (int32_t)startNode;
		h0.minScoreNodeOffset = (int32_t)startOffset;
		h0.maxExactEndposScore = 0;
		h0.maxExactEndposNode = (int32_t)startNode;
		h0.bandwidth = 1;
		h0.scoresNotValid = 0;
		h0.itemStart = 0;
		h0.itemCount = 1;
		h0.pad = 0;
		h0.cells = 0;
		hdr[0] = h0;
		nItems = 1;
		kept = 1;
		buildSliceTable(0);
		if (status != EXT_OK) return;
		int32_t bestX = 0;
		uint32_t numSlices = (len + 63) / 64;
		if (numSlices + 1 > caps.maxSlices)
		{
			fail(EXT_OVERFLOW, 1);
			return;
		}
		for (uint32_t s = 0; s < numSlices && status == EXT_OK; s++)
		{
			const int32_t j = (int32_t)(s * 64);
			const int32_t b = bandLen == 0 ? p.bandwidth : bands[s < bandLen ? s : bandLen - 1];
			const SliceHdr P = hdr[kept - 1];
			const int32_t prevQuit = P.minScore + P.bandwidth;
			Eq4 eq;
			if (sliceEq((uint32_t)j, eq))
			{
				fail(EXT_ASSERT);
				return;
			}
			const uint32_t sliceItemStart = nItems;
			nSlots = 0;
			slotStamp++;
			nEx = 0;
			nPend = 0;
			seqCounter = 0;
			for (uint32_t k = 0; k < P.itemCount; k++)
			{
				Item pi = prevItems[k];
				if (j != 0)
				{
					if (pi.minScore > prevQuit) continue;
					if (g.rec[pi.node].linearizable)
					{
						uint32_t nb = g.rec[pi.node].firstIn;
						int32_t q = findPrev(nb);
						if (q >= 0)
						{
							const Item& qi = prevItems[q];
							if (qi.eEnd < prevQuit && qi.minScore < prevQuit) continue;
						}
					}
				}
				queueInsert(g.rec[pi.node].comp, pi.minScore, pi.node, sourceSlice(pi.sEnd), true);
			}
			if (status != EXT_OK) return;
			if (nPend == 0)
			{
				fail(EXT_ASSERT);
				return;
			}
			int32_t curMin = SMAX - b - 1;
			int32_t rMin = SMAX - b - 1, rMinNode = -1, rMinOff = -1, rMaxX = SMIN, rMaxXNode = 0;
			int64_t cells = 0;
			uint32_t popsThisSlice = 0;
			while (nPend > 0)
			{
				// hard bound on queue pops per slice: every loop of the kernel terminates
				if (++popsThisSlice > kMaxPopsPerSlice)
				{
					fail(EXT_OVERFLOW, 7);
					return;
				}
				uint32_t pi = lpArgMin(pendKey, pendSeq, nPend);
				uint32_t slot = pendSlot[pi];
				uint32_t node = slotNode[slot];
				int32_t itIdx = slotItem[slot];
				bool exists = itIdx >= 0;
				Item it;
				if (!exists)
				{
					if (nItems >= caps.maxItems)
					{
						fail(EXT_OVERFLOW, 5);
						return;
					}
					itIdx = (int32_t)nItems++;
					slotItem[slot] = itIdx;
					itemNode[itIdx] = node;
					it.node = node;
					it.minScore = SMAX;
					it.sVP = it.sVN = it.eVP = it.eVN = 0;
					it.sEnd = it.eEnd = SMAX;
					it.HP = it.HN = 0;
				}
				else
				{
					it = curItems[itIdx - sliceItemStart];
				}
				Slice oldEnd = exists ? mkslice(it.eVP, it.eVN, it.eEnd) : mkslice(0, 0, SMAX);
				int32_t pq = findPrev(node);
				Item prevIt = pq >= 0 ? prevItems[pq] : defaultPrev();
				bool prevHasIn = false;
				if (exists && g.rec[node].inCount == 1) prevHasIn = findPrev(g.rec[node].firstIn) >= 0;
				NodeResult nc = calcNode<false>(node, it, exists, eq, prevIt, pq >= 0, slotHead[slot], mkslice(0, 0, 0), j, prevHasIn, nullptr, rMaxX);
				nodeCalcs++;
				if (status != EXT_OK) return;
				// pop: the node leaves the queue and its queued edges are dropped
				nPend--;
				pendKey[pi] = pendKey[nPend];
				pendSeq[pi] = pendSeq[nPend];
				pendSlot[pi] = pendSlot[nPend];
				slotActive[slot] = 0;
				slotHead[slot] = NIL;
				slotTail[slot] = NIL;
				if (nc.maxX == SMIN || nc.minScore > j + b + 128)
				{
					fail(EXT_ASSERT);
					return;
				}
				curMin = imin(curMin, nc.minScore);
				if (nc.minScore < it.minScore) it.minScore = nc.minScore;
				curItems[itIdx - sliceItemStart] = it;
				Slice newEnd = mkslice(it.eVP, it.eVN, it.eEnd);
				if (!same(newEnd, oldEnd))
				{
					int32_t m = changedMinScore(newEnd, oldEnd);
					if (m == SMAX)
					{
						fail(EXT_ASSERT);
						return;
					}
					if (m <= curMin + b)
					{
						const uint32_t ob = g.rec[node].outBegin, oe = ob + g.rec[node].outCount;
						for (uint32_t e = ob; e < oe; e++)
						{
							const OutEdge edge = g.outEdges[e];
							queueInsert(edge.comp, m, edge.target, newEnd, false);
						}
						if (status != EXT_OK) return;
					}
				}
				if (nc.minScore < rMin)
				{
					rMin = nc.minScore;
					rMinNode = (int32_t)node;
					rMinOff = nc.minScoreNodeOffset;
				}
				if (nc.maxX > rMaxX)
				{
					rMaxX = nc.maxX;
					rMaxXNode = (int32_t)node;
				}
				cells += nc.cells;
				if (cells > p.maxCellsPerSlice) break;
			}
			// fillDPSlice / getXdropSlices checks (src/GraphAlignerBitvectorBanded.h:543-549,633-638)
			double jd = (double)j + 64.0;
			if (rMin > j + 64 || rMaxX == SMIN || (double)rMaxX < -jd * p.errorCost || (double)rMaxX > jd * p.errorCost || rMin < P.minScore || rMinNode < 0)
			{
				fail(EXT_ASSERT);
				return;
			}
			if (rMaxX > bestX) bestX = rMaxX;
			SliceHdr hn;
			hn.minScore = rMin;
			hn.minScoreNode = rMinNode;
			hn.minScoreNodeOffset = rMinOff;
			hn.maxExactEndposScore = rMaxX;
			hn.maxExactEndposNode = rMaxXNode;
			hn.bandwidth = b;
			hn.scoresNotValid = cells >= p.maxCellsPerSlice ? 1 : 0;
			hn.itemStart = sliceItemStart;
			hn.itemCount = nItems - sliceItemStart;
			hn.pad = 0;
			hn.cells = cells;
			if ((double)rMaxX < (double)bestX - (double)p.xdrop)
			{
				nItems = sliceItemStart;
				break;
			}
			for (uint32_t k = 0; k < hn.itemCount; k++) storeItem(sliceItemStart + k, curItems[k]);
			hdr[kept] = hn;
			kept++;
			buildSliceTable(kept - 1);
			if (hn.itemCount > maxBand) maxBand = hn.itemCount;
			Item* t = prevItems;
			prevItems = curItems;
			curItems = t;
			indexPrev(hn.itemCount);
		}
	}

	// --- traceback ---
	GA_HD void pushTrace(const Pos& pos, bool nodeSwitch)
	{
		if (ntrace >= traceCap)
		{
			fail(EXT_OVERFLOW, 6);
			return;
		}
		trace[ntrace++] = packTrace(pos.node, pos.seq, pos.off, nodeSwitch ? 1 : 0);
		back = pos;
	}
	GA_HD Pos tracePos(uint32_t i) const
	{
		uint64_t v = trace[i];
		uint32_t hi = (uint32_t)(v >> 32);
		return mkpos((uint32_t)v, (hi >> 1) & 63, (int32_t)(hi >> 7) - 1);
	}
	// src/GraphAlignerBitvectorCommon.h:760-768
	GA_COLD void checkCircularity()
	{
		for (uint32_t i = ntrace - 1; i-- > 0;)
		{
			Pos q = tracePos(i);
			if (samePos(q, back))
			{
				fail(EXT_ASSERT);
				return;
			}
			if (q.seq != back.seq) return;
		}
	}

	struct Pick
	{
		Pos pos;
		bool sw;
	};
	GA_HD static Pick mkpick(const Pos& p, bool sw)
	{
		Pick r;
		r.pos = p;
		r.sw = sw;
		return r;
	}
	GA_HD int32_t nodeLast(uint32_t n) const { return (int32_t)nodeLen(n) - 1; }

	// pickBacktraceCorner (src/GraphAlignerBitvectorCommon.h:936-1054)
	GA_HD Pick corner(uint32_t cur, uint32_t node, int32_t j, int32_t quit, bool notValid, int32_t prevQuit, bool prevNotValid)
	{
		uint32_t prv = cur - 1;
		int32_t ci = findInSlice(cur, node);
		Item cn = loadItem(ci);
		int32_t scoreHere = valueAt(mkslice(cn.sVP, cn.sVN, cn.sEnd), 0);
		int32_t pn = findInSlice(prv, node);
		uint32_t inB = g.rec[node].inBegin, inE = inB + g.rec[node].inCount;
		if (notValid || scoreHere > quit)
		{
			int32_t smallest = scoreHere + 1;
			Pos sp = mkpos(NIL, NIL, SMIN);
			bool sw = false;
			if (pn >= 0)
			{
				smallest = loadItem(pn).sEnd;
				sp = mkpos(node, 0, j - 1);
				sw = false;
			}
			for (uint32_t e = inB; e < inE; e++)
			{
				uint32_t nb = g.inAdj[e];
				int32_t q = findInSlice(prv, nb);
				if (q >= 0)
				{
					int32_t v = loadItem(q).eEnd;
					if (v <= smallest)
					{
						smallest = v;
						sp = mkpos(nb, nodeLast(nb), j - 1);
						sw = true;
					}
				}
				int32_t c = findInSlice(cur, nb);
				if (c >= 0 && nb != node)
				{
					Item ci2 = loadItem(c);
					int32_t v = valueAt(mkslice(ci2.eVP, ci2.eVN, ci2.eEnd), 0);
					if (v < smallest)
					{
						smallest = v;
						sp = mkpos(nb, nodeLast(nb), j);
						sw = true;
					}
				}
			}
			if (sp.node == NIL) fail(EXT_ASSERT);
			return mkpick(sp, sw);
		}
		bool eq = charMatch(j, node, 0);
		Pos bestInvalid = mkpos(NIL, NIL, SMIN);
		int32_t bestInvalidScore = scoreHere + 1;
		bool bestInvalidSw = false;
		if (pn >= 0)
		{
			int32_t ps = loadItem(pn).sEnd;
			if (ps < scoreHere - 1) fail(EXT_ASSERT);
			if (ps == scoreHere - 1) return mkpick(mkpos(node, 0, j - 1), false);
			if (prevNotValid || ps > prevQuit)
			{
				if (ps < bestInvalidScore)
				{
					bestInvalidScore = ps;
					bestInvalid = mkpos(node, 0, j - 1);
					bestInvalidSw = false;
				}
			}
		}
		for (uint32_t e = inB; e < inE; e++)
		{
			uint32_t nb = g.inAdj[e];
			int32_t c = findInSlice(cur, nb);
			if (c >= 0)
			{
				Item ci2 = loadItem(c);
				int32_t v = valueAt(mkslice(ci2.eVP, ci2.eVN, ci2.eEnd), 0);
				if (v < scoreHere - 1) fail(EXT_ASSERT);
				if (v == scoreHere - 1) return mkpick(mkpos(nb, nodeLast(nb), j), true);
			}
			int32_t q = findInSlice(prv, nb);
			if (q >= 0)
			{
				int32_t cornerScore = loadItem(q).eEnd;
				if (prevNotValid || cornerScore > prevQuit)
				{
					if (cornerScore < bestInvalidScore)
					{
						bestInvalidScore = cornerScore;
						bestInvalid = mkpos(nb, nodeLast(nb), j - 1);
						bestInvalidSw = true;
					}
				}
				else
				{
					if (cornerScore < scoreHere - (eq ? 0 : 1)) fail(EXT_ASSERT);
					if (cornerScore == scoreHere - (eq ? 0 : 1)) return mkpick(mkpos(nb, nodeLast(nb), j - 1), true);
				}
			}
		}
		if (bestInvalidScore < scoreHere + 1) return mkpick(bestInvalid, bestInvalidSw);
		fail(EXT_ASSERT);
		return mkpick(mkpos(0, 0, 0), false);
	}

	// pickBacktraceHorizontalCrossing (src/GraphAlignerBitvectorCommon.h:817-884)
	GA_HD void horizontal(uint32_t cur, uint32_t node, Pos pos, int32_t j, int32_t quit, bool notValid, int32_t prevQuit, bool prevNotValid, Pick& first, Pick& second)
	{
		Item cn = loadItem(findInSlice(cur, node));
		Slice startSlice = mkslice(cn.sVP, cn.sVN, cn.sEnd);
		while ((pos.seq & 63) != 0 && ((startSlice.VP >> (pos.seq & 63)) & 1)) pos.seq--;
		int32_t offset = pos.seq & 63;
		first = mkpick(pos, false);
		if (offset == 0)
		{
			second = corner(cur, node, j, quit, notValid, prevQuit, prevNotValid);
			return;
		}
		bool eq = charMatch(pos.seq, pos.node, pos.off);
		int32_t scoreHere = valueAt(startSlice, offset);
		uint32_t inB = g.rec[node].inBegin, inE = inB + g.rec[node].inCount;
		if (notValid || scoreHere > quit)
		{
			int32_t smallest = valueAt(startSlice, offset - 1);
			Pos sp = mkpos(node, 0, pos.seq - 1);
			bool sw = false;
			for (uint32_t e = inB; e < inE; e++)
			{
				uint32_t nb = g.inAdj[e];
				int32_t c = findInSlice(cur, nb);
				if (c < 0) continue;
				Item ni = loadItem(c);
				Slice ns = mkslice(ni.eVP, ni.eVN, ni.eEnd);
				if (valueAt(ns, offset - 1) <= smallest)
				{
					smallest = valueAt(ns, offset - 1);
					sp = mkpos(nb, nodeLast(nb), pos.seq - 1);
					sw = true;
				}
				if (valueAt(ns, offset) < smallest && nb != node)
				{
					smallest = valueAt(ns, offset);
					sp = mkpos(nb, nodeLast(nb), pos.seq);
					sw = true;
				}
			}
			if (samePos(sp, pos)) fail(EXT_ASSERT);
			second = mkpick(sp, sw);
			return;
		}
		for (uint32_t e = inB; e < inE; e++)
		{
			uint32_t nb = g.inAdj[e];
			int32_t c = findInSlice(cur, nb);
			if (c < 0) continue;
			Item ni = loadItem(c);
			Slice ns = mkslice(ni.eVP, ni.eVN, ni.eEnd);
			if (valueAt(ns, offset) < scoreHere - 1 || valueAt(ns, offset - 1) < scoreHere - (eq ? 0 : 1)) fail(EXT_ASSERT);
			if (valueAt(ns, offset) == scoreHere - 1)
			{
				second = mkpick(mkpos(nb, nodeLast(nb), pos.seq), true);
				return;
			}
			if (valueAt(ns, offset - 1) == scoreHere - (eq ? 0 : 1))
			{
				second = mkpick(mkpos(nb, nodeLast(nb), pos.seq - 1), true);
				return;
			}
		}
		fail(EXT_ASSERT);
		second = mkpick(pos, false);
	}

	// pickBacktraceVerticalCrossing (src/GraphAlignerBitvectorCommon.h:886-934)
	GA_HD void vertical(uint32_t cur, uint32_t node, Pos pos, int32_t j, int32_t quit, bool notValid, int32_t prevQuit, bool prevNotValid, const ColumnStore& cols, Pick& first, Pick& second)
	{
		while (pos.off > 0 && valueAt(colGet(cols, pos.off - 1), 0) == valueAt(colGet(cols, pos.off), 0) - 1) pos.off--;
		first = mkpick(pos, false);
		if (pos.off == 0)
		{
			second = corner(cur, node, j, quit, notValid, prevQuit, prevNotValid);
			return;
		}
		int32_t pn = findInSlice(cur - 1, node);
		if (pn < 0)
		{
			fail(EXT_ASSERT);
			return;
		}
		bool eq = charMatch(pos.seq, pos.node, pos.off);
		Item prevNode = loadItem(pn);
		int32_t scoreHere = valueAt(colGet(cols, pos.off), 0);
		int32_t scoreDiagonal = prevNode.sEnd;
		for (uint32_t i = 1; i + 1 <= pos.off; i++) scoreDiagonal += (int32_t)((prevNode.HP >> i) & 1) - (int32_t)((prevNode.HN >> i) & 1);
		int32_t scoreUp = scoreDiagonal + (int32_t)((prevNode.HP >> pos.off) & 1) - (int32_t)((prevNode.HN >> pos.off) & 1);
		if (prevNotValid || notValid || scoreHere > quit || scoreDiagonal > prevQuit || scoreUp > prevQuit)
		{
			if (scoreDiagonal < scoreUp) second = mkpick(mkpos(pos.node, pos.off - 1, pos.seq - 1), false);
			else second = mkpick(mkpos(pos.node, pos.off, pos.seq - 1), false);
			return;
		}
		if (scoreUp < scoreHere - 1 || scoreDiagonal < scoreHere - (eq ? 0 : 1)) fail(EXT_ASSERT);
		if (scoreUp == scoreHere - 1)
		{
			second = mkpick(mkpos(pos.node, pos.off, pos.seq - 1), false);
			return;
		}
		if (scoreDiagonal == scoreHere - (eq ? 0 : 1))
		{
			second = mkpick(mkpos(pos.node, pos.off - 1, pos.seq - 1), false);
			return;
		}
		fail(EXT_ASSERT);
		second = mkpick(pos, false);
	}

	// recalcNodeWordslice (src/GraphAlignerBitvectorCommon.h:1078-1102). With
	// lastCol, only columns 0..lastCol (those a backtrace entering the node at
	// column lastCol reads); the node's end slice is then not compared with the
	// stored one (:1097-1099) and the end-versus-previous-slice check of the
	// column loop is not made. Both checks compare a recomputation with values
	// the forward pass computed the same way, so they cannot fail without a
	// fault in this code; the host test build with GA_FULL_RECALC_CHECKS
	// recomputes every column and keeps both, and its GAF and read statuses are
	// compared with this build's.
	GA_HD void recalc(uint32_t sliceIdx, uint32_t node, const Eq4& eq, ColumnStore& cols, uint32_t lastCol = NIL)
	{
#ifdef GA_FULL_RECALC_CHECKS
		lastCol = NIL;
#endif
		int32_t ci = findInSlice(sliceIdx, node);
		if (ci < 0)
		{
			fail(EXT_ASSERT);
			return;
		}
		Item stored = loadItem(ci);
		int32_t pq = findInSlice(sliceIdx - 1, node);
		Item prevIt = pq >= 0 ? loadItem(pq) : defaultPrev();
		Item copy = stored;
		const uint32_t cols1 = lastCol == NIL ? NIL : lastCol + 1;
		calcNode<true>(node, copy, true, eq, prevIt, pq >= 0, NIL, mkslice(stored.sVP, stored.sVN, stored.sEnd), (int32_t)((sliceIdx - 1) * 64), false, &cols, SMIN, cols1);
		if (copy.sVP != stored.sVP || copy.sVN != stored.sVN || copy.sEnd != stored.sEnd) fail(EXT_ASSERT);
		if (cols1 >= nodeLen(node) && (copy.eVP != stored.eVP || copy.eVN != stored.eVN || copy.eEnd != stored.eEnd)) fail(EXT_ASSERT);
	}

	// getReverseTraceFromTableExactEndPos + getReverseTraceFromTable
	// (src/GraphAlignerBitvectorCommon.h:496-758), writing trace items end first.
	GA_COLD int32_t traceback()
	{
		uint32_t bestIndex = 1;
		for (uint32_t i = 1; i < kept; i++)
			if (hdr[i].maxExactEndposScore > hdr[bestIndex].maxExactEndposScore) bestIndex = i;
		uint32_t node = (uint32_t)hdr[bestIndex].maxExactEndposNode;
		int32_t score = hdr[bestIndex].maxExactEndposScore;
		int32_t j = (int32_t)((bestIndex - 1) * 64);
		Eq4 eq;
		sliceEq((uint32_t)j, eq);
		ColumnStore cols;
		bindColumns(cols, colWords, colEnd, colsInLds);
		recalc(bestIndex, node, eq, cols);
		if (status != EXT_OK) return 0;
		int32_t cells = imin(64, (int32_t)len - j);
		int32_t nodeOffset = -1, bvOffset = -1;
		uint32_t L = nodeLen(node);
		for (uint32_t i = 0; i < L; i++)
		{
			Slice c = colGet(cols, i);
			int32_t maxScore = lpMaxX(c, j, p.errorCost, maxXCandidates(c, cells, ONES));
			if (maxScore > score)
			{
				fail(EXT_ASSERT);
				return 0;
			}
			if (maxScore == score)
			{
				for (int32_t off = 63; off >= 0; off--)
				{
					if (j + off >= (int32_t)len) continue;
					int32_t xh = (j + off + 1) * 100 - valueAt(c, off) * p.errorCost;
					if (xh > score)
					{
						fail(EXT_ASSERT);
						return 0;
					}
					if (xh == score && (nodeOffset < 0 || off > bvOffset))
					{
						nodeOffset = (int32_t)i;
						bvOffset = off;
					}
				}
			}
		}
		if (nodeOffset < 0)
		{
			fail(EXT_ASSERT);
			return 0;
		}
		int32_t startScore = valueAt(colGet(cols, nodeOffset), bvOffset);
		Pos startPos = mkpos(node, (uint32_t)nodeOffset, j + bvOffset);

		// getReverseTraceFromTable
		int32_t resultScore = startScore;
		ntrace = 0;
		pushTrace(startPos, false);
		uint32_t currentSlice = (uint32_t)(startPos.seq / 64 + 1);
		uint32_t currentNode = NIL;
		sliceEq((currentSlice - 1) * 64, eq);
		int32_t lastScore = SMAX - 1;
		uint32_t steps = 0;
		uint32_t colsUpTo = 0; // last column of the current node's recomputation
		while (back.seq != -1 && status == EXT_OK)
		{
			if (++steps > traceCap)
			{
				fail(EXT_OVERFLOW, 8);
				break;
			}
			uint32_t newSlice = (uint32_t)(back.seq / 64 + 1);
			uint32_t newNode = back.node;
			// the node's columns are recomputed when the walk enters another node
			// or slice, and again when it re-enters the same node of the same
			// slice right of the columns computed so far (through an edge from
			// the node to itself: its first column's in-neighbour is its own last
			// column)
			if (newSlice != currentSlice || newNode != currentNode || back.off > colsUpTo)
			{
				if (newSlice != currentSlice) sliceEq((newSlice - 1) * 64, eq);
				currentSlice = newSlice;
				currentNode = newNode;
				if (currentSlice == 0)
				{
					fail(EXT_ASSERT);
					break;
				}
				recalc(currentSlice, currentNode, eq, cols, back.off);
				colsUpTo = back.off;
				if (status != EXT_OK) break;
			}
			const SliceHdr& ch = hdr[currentSlice];
			const SliceHdr& ph = hdr[currentSlice - 1];
			int32_t cj = (int32_t)((currentSlice - 1) * 64);
			int32_t quit = ch.minScore + ch.bandwidth;
			int32_t prevQuit = ph.minScore + ph.bandwidth;
			bool nv = ch.scoresNotValid != 0, pnv = ph.scoresNotValid != 0;
			Pos bk = back;
			lastScore = valueAt(colGet(cols, bk.off), bk.seq & 63);
			if ((bk.seq & 63) == 0 && bk.off == 0)
			{
				Pick bt = corner(currentSlice, currentNode, cj, quit, nv, prevQuit, pnv);
				if (status != EXT_OK) break;
				pushTrace(bt.pos, bt.sw);
				checkCircularity();
				continue;
			}
			if ((bk.seq & 63) == 0)
			{
				if (findInSlice(currentSlice - 1, currentNode) < 0)
				{
					for (uint32_t i = bk.off - 1; i > 0; i--) pushTrace(mkpos(currentNode, i, back.seq), false);
					pushTrace(mkpos(currentNode, 0, back.seq), false);
					continue;
				}
				Pick first, second;
				vertical(currentSlice, currentNode, bk, cj, quit, nv, prevQuit, pnv, cols, first, second);
				if (status != EXT_OK) break;
				if (first.pos.node != bk.node || first.pos.seq != bk.seq || first.pos.off > bk.off)
				{
					fail(EXT_ASSERT);
					break;
				}
				if (first.pos.off != bk.off)
					for (uint32_t o = bk.off - 1; o != first.pos.off; o--) pushTrace(mkpos(first.pos.node, o, first.pos.seq), false);
				if (!samePos(first.pos, back)) pushTrace(first.pos, first.sw);
				if (samePos(second.pos, back))
				{
					fail(EXT_ASSERT);
					break;
				}
				pushTrace(second.pos, second.sw);
				continue;
			}
			if (bk.off == 0)
			{
				Pick first, second;
				horizontal(currentSlice, currentNode, bk, cj, quit, nv, prevQuit, pnv, first, second);
				if (status != EXT_OK) break;
				if (first.pos.node != bk.node || first.pos.off != bk.off || first.pos.seq > bk.seq)
				{
					fail(EXT_ASSERT);
					break;
				}
				if (first.pos.seq != bk.seq)
					for (int32_t sp = bk.seq - 1; sp != first.pos.seq; sp--) pushTrace(mkpos(first.pos.node, first.pos.off, sp), false);
				if (!samePos(first.pos, back)) pushTrace(first.pos, first.sw);
				if (samePos(second.pos, back))
				{
					fail(EXT_ASSERT);
					break;
				}
				pushTrace(second.pos, second.sw);
				checkCircularity();
				continue;
			}
			// pickBacktraceInside (src/GraphAlignerBitvectorCommon.h:770-815) on the
			// columns' delta bits. With here the score of cell (c, r), vd_c(r) the
			// column's vertical delta (VP/VN bit) and hd_c(r) the horizontal delta of
			// the step from column c - 1 (Ph/Mh bit):
			//   vertical   = here - vd_c(r)
			//   horizontal = here - hd_c(r)
			//   diagonal   = here - hd_c(r) - vd_{c-1}(r)
			// so every choice of the walk is a test of a few bits at row r.
			uint32_t hori = bk.off;
			int32_t vert = bk.seq - cj;
			if (hori > 0 && vert > 0)
			{
				Slice cc = colGet(cols, hori);
				Slice cp = colGet(cols, hori - 1);
				ColStep st = colGetStep(cols, hori);
				while (status == EXT_OK)
				{
					const uint64_t bit = (uint64_t)1 << vert;
					const int32_t hd = (int32_t)((st.ph & bit) != 0) - (int32_t)((st.mh & bit) != 0);
					const int32_t hdv = hd + (int32_t)((cp.VP & bit) != 0) - (int32_t)((cp.VN & bit) != 0);
					// diagonal == here - (characterMatch ? 0 : 1) when hdv == want
					const int32_t want = (st.eq & bit) ? 0 : 1;
					// vertical and horizontal never score below here - 1 (|delta| <= 1)
					if (hdv > want)
					{
						fail(EXT_ASSERT);
						break;
					}
					if (cc.VP & bit)
					{
						vert--;
						pushTrace(mkpos(bk.node, hori, vert + cj), false);
						if (vert == 0) break;
						continue;
					}
					if (hdv == want) vert--;
					else if (hd != 1)
					{
						fail(EXT_ASSERT);
						break;
					}
					hori--;
					pushTrace(mkpos(bk.node, hori, vert + cj), false);
					if (hori == 0 || vert == 0) break;
					cc = cp;
					cp = colGet(cols, hori - 1);
					st = colGetStep(cols, hori);
				}
			}
		}
		if (status != EXT_OK) return 0;
		if (resultScore < lastScore)
		{
			fail(EXT_ASSERT);
			return 0;
		}
		resultScore -= lastScore;
		// final walk inside the initial slice (src/GraphAlignerBitvectorCommon.h:721-756)
		{
			int32_t i0 = findInSlice(0, back.node);
			if (i0 < 0)
			{
				fail(EXT_ASSERT);
				return 0;
			}
			Item n0 = loadItem(i0);
			uint32_t L0 = nodeLen(back.node);
			// before-slice score at column c = sEnd + prefix(HP) - prefix(HN) over bits 1..c
			auto beforeScore = [&](uint32_t c) -> int32_t {
				uint64_t m = c >= 63 ? ONES : ((((uint64_t)1) << (c + 1)) - 1);
				m &= ~(uint64_t)1;
				return n0.sEnd + popc(n0.HP & m) - popc(n0.HN & m);
			};
			if (beforeScore(L0 - 1) != n0.eEnd)
			{
				fail(EXT_ASSERT);
				return 0;
			}
			while (beforeScore(back.off) != 0 && back.off > 0 && beforeScore(back.off - 1) == beforeScore(back.off) - 1)
			{
				pushTrace(mkpos(back.node, back.off - 1, back.seq), false);
				if (status != EXT_OK) return 0;
			}
			if (back.off == 0 && beforeScore(0) != 0)
			{
				for (uint32_t e = g.rec[back.node].inBegin, inE = e + g.rec[back.node].inCount; e < inE; e++)
				{
					uint32_t nb = g.inAdj[e];
					int32_t q = findInSlice(0, nb);
					if (q < 0) continue;
					Item qi = loadItem(q);
					if (before(mkslice(qi.eVP, qi.eVN, qi.eEnd)) == beforeScore(0) - 1)
					{
						pushTrace(mkpos(nb, nodeLast(nb), back.seq), true);
						break;
					}
				}
			}
		}
		return resultScore;
	}

	GA_HD void run(ExtResult& out)
	{
		status = EXT_OK;
		overflowWhat = 0;
		wordSteps = 0;
		nodeCalcs = 0;
		maxBand = 0;
		ntrace = 0;
		nItems = kept = nSlots = nEx = nPend = seqCounter = 0;
		sliceTabUsed = 0;
		slotStamp = 0;
		for (uint32_t i = 0; i <= slotTabMask; i++) slotTabKey[i] = 0;
		back = mkpos(NIL, 0, 0);
		forward();
		if (summaries)
		{
			for (uint32_t i = 1; i < kept; i++)
			{
				const SliceHdr& h = hdr[i];
				SliceSummary ss;
				ss.j = (int64_t)(i - 1) * 64;
				ss.minScore = h.minScore;
				ss.pad0 = 0;
				ss.minScoreNode = h.minScoreNode;
				ss.minScoreNodeOffset = h.minScoreNodeOffset;
				ss.maxExactEndposScore = h.maxExactEndposScore;
				ss.pad1 = 0;
				ss.maxExactEndposNode = h.maxExactEndposNode;
				ss.cellsProcessed = h.cells;
				ss.numNodes = h.itemCount;
				ss.itemsHash = lpDigest(items, h.itemStart, h.itemCount);
				ss.scoresNotValid = h.scoresNotValid;
				ss.pad2 = 0;
				summaries[i - 1] = ss;
			}
		}
		int32_t score = SMAX;
		if (status == EXT_OK)
		{
			if (kept <= 1) status = EXT_FAILED;
			else if (hdr[kept - 1].minScore < 0 || hdr[kept - 1].minScore > (int32_t)len + 128) fail(EXT_ASSERT);
			else score = traceback();
		}
		out.status = status;
		out.score = status == EXT_OK ? score : SMAX;
		out.ntrace = status == EXT_OK ? ntrace : 0;
		out.nslices = kept > 0 ? kept - 1 : 0;
		out.wordSteps = wordSteps;
		out.nodeCalcs = nodeCalcs;
		out.maxBand = maxBand;
		out.overflowWhat = overflowWhat;
		out.clockStart = out.clockEnd = 0;
	}
};

GA_HHD uint64_t alignUp(uint64_t x) { return (x + 255) & ~(uint64_t)255; }
// Entries of the per-slice node tables: each slice's table has at most
// 4 * itemCount entries and at least 4.
GA_HHD uint64_t sliceTabEntries(const ArenaCaps& c) { return 4 * (uint64_t)c.maxItems + 4 * (uint64_t)c.maxSlices; }
// Entries of the slot table: a power of two at least twice the slot budget.
GA_HHD uint32_t slotTabSize(const ArenaCaps& c)
{
	uint32_t t = 4;
	while (t < 2 * c.maxSlots) t <<= 1;
	return t;
}

// Bytes of one wave's workspace for the given capacities.
GA_HHD uint64_t arenaBytes(const ArenaCaps& c)
{
	uint64_t b = 0;
	b += alignUp((uint64_t)c.maxSlices * sizeof(SliceHdr));
	b += alignUp((uint64_t)c.maxItems * sizeof(Item));
	b += alignUp((uint64_t)c.maxItems * 4);
	b += 5 * alignUp((uint64_t)c.maxSlots * 4);
	b += alignUp(sliceTabEntries(c) * 8);
	b += alignUp((uint64_t)slotTabSize(c) * 8) + alignUp((uint64_t)slotTabSize(c) * 4);
	b += 2 * alignUp((uint64_t)c.maxExtras * 8) + 2 * alignUp((uint64_t)c.maxExtras * 4);
	b += alignUp((uint64_t)c.maxPending * 8) + 2 * alignUp((uint64_t)c.maxPending * 4);
	b += 2 * alignUp((uint64_t)c.maxSlots * sizeof(Item)) + alignUp((uint64_t)slotTabSize(c) * 8);
	b += alignUp(kColumnWords * 8);
	return b;
}

GA_HD void carveArena(Extension& e, unsigned char* base, const ArenaCaps& c)
{
	e.caps = c;
	uint64_t o = 0;
	auto take = [&](uint64_t bytes) {
		unsigned char* p = base + o;
		o += alignUp(bytes);
		return p;
	};
	e.hdr = (SliceHdr*)take((uint64_t)c.maxSlices * sizeof(SliceHdr));
	e.items = (Item*)take((uint64_t)c.maxItems * sizeof(Item));
	e.itemNode = (uint32_t*)take((uint64_t)c.maxItems * 4);
	e.slotNode = (uint32_t*)take((uint64_t)c.maxSlots * 4);
	e.slotItem = (int32_t*)take((uint64_t)c.maxSlots * 4);
	e.slotHead = (uint32_t*)take((uint64_t)c.maxSlots * 4);
	e.slotTail = (uint32_t*)take((uint64_t)c.maxSlots * 4);
	e.slotActive = (uint32_t*)take((uint64_t)c.maxSlots * 4);
	e.sliceTab = (uint64_t*)take(sliceTabEntries(c) * 8);
	e.sliceTabCap = (uint32_t)sliceTabEntries(c);
	e.slotTabKey = (uint64_t*)take((uint64_t)slotTabSize(c) * 8);
	e.slotTabVal = (uint32_t*)take((uint64_t)slotTabSize(c) * 4);
	e.slotTabMask = slotTabSize(c) - 1;
	e.exVP = (uint64_t*)take((uint64_t)c.maxExtras * 8);
	e.exVN = (uint64_t*)take((uint64_t)c.maxExtras * 8);
	e.exEnd = (int32_t*)take((uint64_t)c.maxExtras * 4);
	e.exNext = (uint32_t*)take((uint64_t)c.maxExtras * 4);
	e.pendKey = (uint64_t*)take((uint64_t)c.maxPending * 8);
	e.pendSeq = (uint32_t*)take((uint64_t)c.maxPending * 4);
	e.pendSlot = (uint32_t*)take((uint64_t)c.maxPending * 4);
	e.curItems = (Item*)take((uint64_t)c.maxSlots * sizeof(Item));
	e.prevItems = (Item*)take((uint64_t)c.maxSlots * sizeof(Item));
	e.prevTab = (uint64_t*)take((uint64_t)slotTabSize(c) * 8);
	e.prevTabMask = slotTabSize(c) - 1;
	e.colWords = (uint64_t*)take(kColumnWords * 8);
	e.colEnd = (int32_t*)(e.colWords + 320);
	e.colsInLds = false;
}

// The per-slice band queue (slots, their node table, queued edges, pending
// nodes) held in LDS by the single-lane kernel when the capacities allow: it is
// rebuilt at every slice, small (the 5 Mbp variation-graph workload peaks at 14
// slots, 29 queued edges and 9 pending nodes in a slice) and on the path of
// every queue pop, where the workspace's HBM round trips cost most. An
// extension that outgrows it fails with EXT_OVERFLOW and is rerun with larger
// capacities and the queue in its workspace.
// (capacities kLdsSlots, kLdsExtras, kLdsPending: ga_device_types.h)
constexpr uint32_t kLdsSlotTab = 2 * kLdsSlots; // slotTabSize of kLdsSlots
// 64-bit words: slot table keys, edge VP and VN words, pending keys, the 32-bit
// arrays two to a word, then the current and previous slice's items (8 words
// each) and the previous slice's node table. The traceback's column store
// (kColumnWords) shares the same LDS: the forward pass and the traceback do not
// overlap.
constexpr uint32_t kQueueWords = kLdsSlotTab + 2 * kLdsExtras + kLdsPending + (kLdsSlotTab + 5 * kLdsSlots + 2 * kLdsExtras + 2 * kLdsPending) / 2 + 2 * kLdsSlots * 8 + kLdsSlotTab;
static_assert(sizeof(Item) == 64, "Item is eight words");
static_assert(kQueueWords >= kColumnWords, "the queue region covers the column store");
GA_HHD bool queueFitsLds(const ArenaCaps& c) { return c.maxSlots <= kLdsSlots && c.maxExtras <= kLdsExtras && c.maxPending <= kLdsPending; }
GA_HD void bindQueue(Extension& e, uint64_t* w)
{
	e.slotTabKey = w;
	w += kLdsSlotTab;
	e.exVP = w;
	w += kLdsExtras;
	e.exVN = w;
	w += kLdsExtras;
	e.pendKey = w;
	w += kLdsPending;
	uint32_t* u = (uint32_t*)w;
	e.slotTabVal = u;
	u += kLdsSlotTab;
	e.slotNode = u;
	u += kLdsSlots;
	e.slotItem = (int32_t*)u;
	u += kLdsSlots;
	e.slotHead = u;
	u += kLdsSlots;
	e.slotTail = u;
	u += kLdsSlots;
	e.slotActive = u;
	u += kLdsSlots;
	e.exEnd = (int32_t*)u;
	u += kLdsExtras;
	e.exNext = u;
	u += kLdsExtras;
	e.pendSeq = u;
	u += kLdsPending;
	e.pendSlot = u;
	u += kLdsPending;
	e.curItems = (Item*)u;
	e.prevItems = e.curItems + kLdsSlots;
	e.prevTab = (uint64_t*)(e.prevItems + kLdsSlots);
	e.prevTabMask = kLdsSlotTab - 1;
}

// Per-request capacities that scale with the read length. caps.maxItems is the
// band-item budget per slice.
GA_HHD uint32_t requestSlices(uint32_t len) { return len / 64 + 3; }
GA_HHD uint32_t requestItems(uint32_t len, const ArenaCaps& caps) { return requestSlices(len) * caps.maxItems + 1024; }

// Runs one request on the calling wave.
// colStore, when given, holds the traceback's column store (kColumnWords
// words) in place of the arena's copy; the backtrace reads it at
// every step. colStoreLds says it is an LDS array (a compile-time constant at
// the kernels' call sites, so only one access path is built); it requires colStore.
// planes, when given, are the match-mask planes of seqbuf (planeBlock).
// With queueLds (a compile-time constant at the call sites) the band queue lives
// in queueStore (kQueueWords words of LDS); the capacities must fit it
// (queueFitsLds).
GA_HD void runRequest(const DevGraph& g, const ExtParams& p, const ExtRequest& r, const char* seqbuf, uint64_t* tracebuf, unsigned char* arena, const ArenaCaps& caps, SliceSummary* summaries, ExtResult& out, uint64_t* colStore = nullptr, const uint64_t* planes = nullptr, bool colStoreLds = false, uint64_t* queueStore = nullptr, bool queueLds = false)
{
	Extension e;
	carveArena(e, arena, caps);
	if (queueLds) bindQueue(e, queueStore);
	if (colStore || colStoreLds)
	{
		e.colWords = colStore;
		e.colEnd = (int32_t*)(colStore + 320);
	}
	e.colsInLds = colStoreLds;
	// only one-thread workgroups keep the column store in LDS
	e.aloneInWave = colStoreLds;
	e.g = g;
	e.p = p;
	e.seq = seqbuf + r.seqOffset;
	e.planes = planes;
	e.seqBase = r.seqOffset;
	e.len = r.len;
	e.startNode = r.startNode;
	e.startOffset = r.startOffset;
	e.bandLen = r.bandLen;
	e.bands = r.bandLen ? p.bandSched + r.bandOff : nullptr;
	e.trace = tracebuf + r.traceOffset;
	e.traceCap = r.traceCap;
	e.summaries = summaries;
	e.run(out);
}

} // namespace gad
