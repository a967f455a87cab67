// Minimizer index construction (see ga_index.hpp).
#include "ga_index.hpp"
#include "ga_seed_core.h"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

namespace ga {

static int baseCode(char c) { return gad::readBaseCode((unsigned char)c); }

namespace {

// Window of (position, k-mer, hash) candidates, a fixed-size ring used as the
// monotone deque of src/MinimizerSeeder.cpp:104-189.
struct Window
{
	struct Entry
	{
		size_t pos;
		uint64_t kmer;
		uint64_t hash;
	};
	std::vector<Entry> buf;
	size_t head = 0, size = 0;
	explicit Window(size_t cap) : buf(cap + 2) {}
	Entry& at(size_t i) { return buf[(head + i) % buf.size()]; }
	Entry& front() { return at(0); }
	Entry& back() { return at(size - 1); }
	void clear() { head = size = 0; }
	void popFront()
	{
		head = (head + 1) % buf.size();
		size--;
	}
	void popBack() { size--; }
	void push(size_t pos, uint64_t kmer, uint64_t hash)
	{
		buf[(head + size) % buf.size()] = Entry{pos, kmer, hash};
		size++;
	}
};

// Reports (position of the last base, k-mer) for every window minimum; all
// positions sharing the minimal hash are reported (the reference's rule).
template <typename F>
void forEachMinimizer(const std::string& s, size_t k, size_t w, F report)
{
	if (s.size() < k) return;
	const size_t span = w - k + 1;
	const uint64_t mask = ~(~0ull << (2 * k));
	Window win(span + 1);
	size_t offset = 0;
	while (true)
	{
		while (offset < s.size() && baseCode(s[offset]) < 0) offset++;
		if (offset + w > s.size()) return;
		uint64_t kmer = 0;
		size_t bad = s.size();
		for (size_t i = 0; i < k; i++)
		{
			int c = baseCode(s[offset + i]);
			if (c < 0)
			{
				bad = i;
				break;
			}
			kmer = (kmer << 2) | (uint64_t)c;
		}
		if (bad != s.size())
		{
			offset += bad;
			continue;
		}
		win.clear();
		win.push(offset + k - 1, kmer, gad::seedKmerHash(kmer));
		bool restart = false;
		for (size_t i = k; i < k + span; i++)
		{
			int c = baseCode(s[offset + i]);
			if (c < 0)
			{
				offset += i;
				restart = true;
				break;
			}
			kmer = ((kmer << 2) & mask) | (uint64_t)c;
			uint64_t h = gad::seedKmerHash(kmer);
			while (win.size > 0 && win.back().hash > h) win.popBack();
			win.push(offset + i, kmer, h);
		}
		if (restart) continue;
		for (size_t i = 0; i < win.size && win.at(i).hash == win.front().hash; i++) report(win.at(i).pos, win.at(i).kmer);
		for (size_t i = k + span; offset + i < s.size(); i++)
		{
			int c = baseCode(s[offset + i]);
			if (c < 0)
			{
				offset += i;
				restart = true;
				break;
			}
			kmer = ((kmer << 2) & mask) | (uint64_t)c;
			uint64_t h = gad::seedKmerHash(kmer);
			uint64_t oldMin = win.front().hash;
			bool frontPopped = false;
			while (win.size > 0 && win.front().pos <= offset + i - span)
			{
				frontPopped = true;
				win.popFront();
			}
			if (frontPopped)
				while (win.size >= 2 && win.front().hash == win.at(1).hash) win.popFront();
			while (win.size > 0 && win.back().hash > h) win.popBack();
			win.push(offset + i, kmer, h);
			if (win.front().hash != oldMin)
			{
				for (size_t q = 0; q < win.size && win.at(q).hash == win.front().hash; q++) report(win.at(q).pos, win.at(q).kmer);
			}
			else if (win.back().hash == win.front().hash)
			{
				report(win.back().pos, win.back().kmer);
			}
		}
		if (!restart) return;
	}
}

} // namespace

const KmerSlot* MinimizerIndex::find(uint64_t kmer) const
{
	if (table.empty()) return nullptr;
	uint64_t i = gad::seedKmerHash(kmer) & mask;
	while (true)
	{
		const KmerSlot& s = table[i];
		if (s.kmer == kmer) return &s;
		if (s.kmer == ~0ull) return nullptr;
		i = (i + 1) & mask;
	}
}

MinimizerIndex buildMinimizerIndex(const GraphHost& g, uint32_t k, uint32_t w, double discardFraction)
{
	if (k < 1 || k > 31 || w < k) throw std::runtime_error("minimizer length must be 1..31 and not above the window size");
	MinimizerIndex idx;
	idx.uid = nextObjectUid();
	idx.k = k;
	idx.w = w;
	// graph minimizers are kept only from the first split piece that has an
	// in-edge from another bigraph node onwards (DBG overlaps, :320-337)
	std::unordered_map<int, uint64_t> minStart;
	for (size_t i = 0; i < g.numNodes(); i++)
	{
		uint64_t& ms = minStart[g.nodeId[i]];
		for (uint32_t e = g.inOff[i]; e < g.inOff[i + 1]; e++)
		{
			if (g.nodeId[g.inAdj[e]] != g.nodeId[i])
			{
				ms = std::max<uint64_t>(ms, g.nodeOff[i]);
				break;
			}
		}
	}
	std::vector<std::pair<uint64_t, uint64_t>> found;
	std::string seq;
	for (int id : g.lookupOrder)
	{
		uint64_t size = g.originalSize.at(id);
		seq.resize(size);
		const auto& pieces = g.lookup.at(id);
		for (uint32_t node : pieces)
			for (uint32_t p = 0; p < g.nodeLen[node]; p++) seq[g.nodeOff[node] + p] = g.nodeChar(node, p);
		uint64_t ms = minStart.at(id);
		forEachMinimizer(seq, k, w, [&](size_t pos, uint64_t kmer) {
			if (pos < ms) return;
			uint32_t node = g.unitigNode(id, pos);
			found.emplace_back(kmer, ((uint64_t)node << 6) + (pos - g.nodeOff[node]));
		});
	}
	// group by k-mer; within a k-mer keep reverse discovery order
	std::unordered_map<uint64_t, uint32_t> count;
	count.reserve(found.size());
	for (auto& f : found) count[f.first]++;
	idx.numKmers = count.size();
	uint64_t cap = 16;
	while (cap < 2 * idx.numKmers + 16) cap <<= 1;
	idx.mask = cap - 1;
	idx.table.assign(cap, KmerSlot{~0ull, 0, 0});
	uint32_t next = 0;
	std::unordered_map<uint64_t, uint32_t> fill;
	fill.reserve(count.size());
	for (auto& f : found)
	{
		if (fill.count(f.first)) continue;
		uint32_t c = count[f.first];
		uint64_t i = gad::seedKmerHash(f.first) & idx.mask;
		while (idx.table[i].kmer != ~0ull) i = (i + 1) & idx.mask;
		idx.table[i] = KmerSlot{f.first, next, c};
		fill[f.first] = next + c;
		next += c;
	}
	idx.positions.resize(found.size());
	for (auto& f : found) idx.positions[--fill[f.first]] = f.second;
	// maxCount (src/MinimizerSeeder.cpp:557-575): the reference leaves out one
	// arbitrary key (the last BBHash slot); here the key with the smallest count.
	std::vector<uint32_t> counts;
	counts.reserve(count.size());
	for (auto& c : count) counts.push_back(c.second);
	std::sort(counts.begin(), counts.end());
	if (!counts.empty()) counts.erase(counts.begin());
	idx.maxCount = 0;
	if (!counts.empty())
	{
		size_t index = counts.size() * (1.0 - discardFraction);
		if (index == counts.size()) index = counts.size() - 1;
		idx.maxCount = (uint64_t)counts[index] + 1;
	}
	return idx;
}

} // namespace ga
