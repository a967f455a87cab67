// Seed-cluster features, outcome records and pruning of the learned proposal.
#include "ga_proposal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <cmath>
#include <memory>
#include <thread>

namespace ga {

namespace {

constexpr uint64_t kNone = ~0ull;

// per graph: size of each split node's strongly connected component and the
// length of each chain (bases up to its last node's end)
struct GraphShape
{
	uint64_t uid = 0;
	std::vector<uint32_t> compSize;
	std::vector<uint64_t> chainLen;
};

std::shared_ptr<const GraphShape> shapeOf(const GraphHost& g)
{
	static std::mutex m;
	static std::vector<std::shared_ptr<const GraphShape>> cache;
	std::lock_guard<std::mutex> lock(m);
	for (const auto& c : cache)
		if (c->uid == g.uid && g.uid != 0) return c;
	auto s = std::make_shared<GraphShape>();
	s->uid = g.uid;
	const size_t n = g.nodeLen.size();
	uint32_t maxComp = 0, maxChain = 0;
	for (size_t i = 0; i < n; i++)
	{
		maxComp = std::max(maxComp, g.comp[i]);
		maxChain = std::max(maxChain, g.chain[i]);
	}
	std::vector<uint32_t> perComp(n ? maxComp + 1 : 0, 0);
	for (size_t i = 0; i < n; i++) perComp[g.comp[i]]++;
	s->compSize.resize(n);
	for (size_t i = 0; i < n; i++) s->compSize[i] = perComp[g.comp[i]];
	s->chainLen.assign(n ? maxChain + 1 : 0, 0);
	for (size_t i = 0; i < n; i++) s->chainLen[g.chain[i]] = std::max(s->chainLen[g.chain[i]], g.chainPos[i] + g.nodeLen[i]);
	if (cache.size() >= 4) cache.erase(cache.begin());
	cache.push_back(s);
	return s;
}

// split node and offset of a seed (seeds from the minimizer seeder carry them)
std::pair<uint32_t, uint64_t> splitPosition(const GraphHost& g, const SeedHit& s)
{
	if (s.alignmentGraphNodeId != kNone) return {(uint32_t)s.alignmentGraphNodeId, s.alignmentGraphNodeOffset};
	const int forwardId = s.reverse ? s.nodeID * 2 + 1 : s.nodeID * 2;
	const uint32_t node = g.unitigNode(forwardId, s.nodeOffset);
	return {node, s.nodeOffset - g.nodeOff[node]};
}

int baseCode(char c)
{
	switch (c)
	{
	case 'A': case 'a': return 0;
	case 'C': case 'c': return 1;
	case 'G': case 'g': return 2;
	case 'T': case 't': case 'U': case 'u': return 3;
	default: return -1;
	}
}

// Shannon entropy (bits) of the k-mers (k = 2 or 3) of seq[from, to)
double kmerEntropy(const char* seq, uint64_t from, uint64_t to, int k)
{
	uint32_t counts[64] = {0};
	uint32_t total = 0;
	for (uint64_t p = from; p + k <= to; p++)
	{
		int code = 0;
		bool ok = true;
		for (int q = 0; q < k && ok; q++)
		{
			const int b = baseCode(seq[p + q]);
			ok = b >= 0;
			code = code * 4 + b;
		}
		if (!ok) continue;
		counts[code]++;
		total++;
	}
	if (total == 0) return 0;
	double h = 0;
	for (int c = 0; c < (1 << (2 * k)); c++)
		if (counts[c])
		{
			const double p = (double)counts[c] / total;
			h -= p * std::log2(p);
		}
	return h;
}

uint32_t longestRun(const char* seq, uint64_t from, uint64_t to)
{
	uint32_t best = 0, run = 0;
	for (uint64_t p = from; p < to; p++)
	{
		run = p > from && seq[p] == seq[p - 1] ? run + 1 : 1;
		best = std::max(best, run);
	}
	return best;
}

float lg(double x) { return (float)std::log2(1.0 + std::max(0.0, x)); }

template <class F>
void forEachRead(size_t n, int threads, F fn)
{
	const size_t t = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, threads), n / 64 + 1));
	if (t == 1)
	{
		for (size_t i = 0; i < n; i++) fn(i);
		return;
	}
	std::vector<std::thread> pool;
	std::atomic<size_t> next{0};
	std::mutex em;
	std::exception_ptr err;
	for (size_t w = 0; w < t; w++)
		pool.emplace_back([&]() {
			try
			{
				for (size_t i = next++; i < n; i = next++) fn(i);
			}
			catch (...)
			{
				std::lock_guard<std::mutex> lock(em);
				if (!err) err = std::current_exception();
			}
		});
	for (auto& th : pool) th.join();
	if (err) std::rethrow_exception(err);
}

} // namespace

uint32_t seedClusters(const GraphHost& g, const std::vector<SeedHit>& seeds, std::vector<uint32_t>& clusterOf)
{
	// the grouping of orderSeedsByChaining: per chain, seeds sorted by diagonal,
	// a new cluster where the diagonal jumps by more than 100
	struct Key
	{
		uint32_t chain;
		uint64_t diag;
		uint32_t seed;
	};
	std::vector<Key> keys(seeds.size());
	for (size_t i = 0; i < seeds.size(); i++)
	{
		const auto pos = splitPosition(g, seeds[i]);
		const uint64_t at = g.chainPos[pos.first] + pos.second;
		keys[i] = Key{g.chain[pos.first], at >= seeds[i].seqPos ? at - seeds[i].seqPos : 0, (uint32_t)i};
	}
	std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) { return a.chain != b.chain ? a.chain < b.chain : a.diag != b.diag ? a.diag < b.diag : a.seed < b.seed; });
	std::vector<uint32_t> group(seeds.size());
	uint32_t groups = 0;
	for (size_t k = 0; k < keys.size(); k++)
	{
		if (k == 0 || keys[k].chain != keys[k - 1].chain || keys[k].diag > keys[k - 1].diag + 100) groups++;
		group[keys[k].seed] = groups - 1;
	}
	// renumber by first appearance in loop order
	std::vector<uint32_t> number(groups, ~0u);
	uint32_t n = 0;
	clusterOf.resize(seeds.size());
	for (size_t i = 0; i < seeds.size(); i++)
	{
		uint32_t& c = number[group[i]];
		if (c == ~0u) c = n++;
		clusterOf[i] = c;
	}
	return n;
}

void clusterFeatures(const GraphHost& g, uint64_t maxCount, const char* seq, uint32_t len, const std::vector<SeedHit>& seeds, const std::vector<uint32_t>& clusterOf, uint32_t nclusters, uint32_t rows, float* out)
{
	rows = std::min(rows, nclusters);
	auto shape = shapeOf(g);
	const double L = std::max<uint32_t>(1, len);
	struct C
	{
		std::vector<uint32_t> members;
		uint64_t goodness = 0, matchBp = 0;
		uint64_t minPos = kNone, maxPos = 0, minDiag = kNone, maxDiag = 0;
		double diagSum = 0, diagSq = 0, rawSum = 0;
		uint64_t minRaw = kNone;
		double inSum = 0, outSum = 0;
		uint32_t inMax = 0, outMax = 0, inMulti = 0, outMulti = 0, inScc = 0;
		uint32_t sccMax = 0;
		uint64_t chainLen = 0;
		bool reverse = false;
		uint64_t firstPos = 0;
		uint64_t spanFrom = 0, spanTo = 0;
	};
	std::vector<C> cl(nclusters);
	for (size_t i = 0; i < seeds.size(); i++) cl[clusterOf[i]].members.push_back((uint32_t)i);
	uint64_t best = 0;
	size_t maxSize = 0;
	for (uint32_t c = 0; c < nclusters; c++)
	{
		C& x = cl[c];
		const SeedHit& first = seeds[x.members[0]];
		x.reverse = first.reverse;
		x.firstPos = first.seqPos;
		std::vector<std::pair<uint64_t, uint64_t>> spans;
		for (uint32_t s : x.members)
		{
			const SeedHit& h = seeds[s];
			x.goodness = std::max(x.goodness, h.seedGoodness);
			x.minPos = std::min(x.minPos, h.seqPos);
			x.maxPos = std::max(x.maxPos, h.seqPos);
			const uint64_t start = h.seqPos + 1 >= h.matchLen ? h.seqPos + 1 - h.matchLen : 0;
			spans.emplace_back(start, h.seqPos + 1);
			const auto pos = splitPosition(g, h);
			const uint64_t at = g.chainPos[pos.first] + pos.second;
			const uint64_t diag = at >= h.seqPos ? at - h.seqPos : 0;
			x.minDiag = std::min(x.minDiag, diag);
			x.maxDiag = std::max(x.maxDiag, diag);
			x.diagSum += (double)diag;
			x.diagSq += (double)diag * (double)diag;
			x.rawSum += (double)h.rawSeedGoodness;
			x.minRaw = std::min(x.minRaw, h.rawSeedGoodness);
			const uint32_t in = g.inOff[pos.first + 1] - g.inOff[pos.first];
			const uint32_t outd = g.outOff[pos.first + 1] - g.outOff[pos.first];
			x.inSum += in;
			x.outSum += outd;
			x.inMax = std::max(x.inMax, in);
			x.outMax = std::max(x.outMax, outd);
			x.inMulti += in > 1;
			x.outMulti += outd > 1;
			const uint32_t scc = shape->compSize[pos.first];
			x.sccMax = std::max(x.sccMax, scc);
			x.inScc += scc > 1;
			x.chainLen = std::max(x.chainLen, shape->chainLen[g.chain[pos.first]]);
		}
		std::sort(spans.begin(), spans.end());
		uint64_t covered = 0, end = 0;
		for (const auto& sp : spans)
		{
			const uint64_t from = std::max(sp.first, end);
			if (sp.second > from) covered += sp.second - from;
			end = std::max(end, sp.second);
		}
		x.matchBp = covered;
		x.spanFrom = spans.front().first;
		x.spanTo = x.maxPos + 1;
		best = std::max(best, x.goodness);
		maxSize = std::max(maxSize, x.members.size());
	}
	// read coverage of the clusters ranked before each one
	// read positions covered by the spans of the clusters ranked before the
	// current one: disjoint intervals in position order
	std::vector<std::pair<uint64_t, uint64_t>> coveredBefore;
	uint64_t coveredCount = 0;
	auto coveredIn = [&](uint64_t from, uint64_t to) {
		uint64_t n = 0;
		for (const auto& iv : coveredBefore)
		{
			const uint64_t lo = std::max(from, iv.first), hi = std::min(to, iv.second);
			if (hi > lo) n += hi - lo;
		}
		return n;
	};
	for (uint32_t c = 0; c < rows; c++)
	{
		const C& x = cl[c];
		float* f = out + (size_t)c * kClusterFeatures;
		std::fill(f, f + kClusterFeatures, 0.0f);
		const double n = (double)x.members.size();
		const double span = (double)(x.spanTo - x.spanFrom);
		uint64_t spanCovered = 0;
		spanCovered = coveredIn(x.spanFrom, std::min<uint64_t>(x.spanTo, len));
		uint32_t overlapping = 0;
		double bestOverlapGoodness = 0;
		uint64_t bestClusterOverlap = 0;
		for (uint32_t d = 0; d < c; d++)
		{
			const C& y = cl[d];
			const uint64_t lo = std::max(x.spanFrom, y.spanFrom), hi = std::min(x.spanTo, y.spanTo);
			if (hi <= lo) continue;
			overlapping++;
			bestOverlapGoodness = std::max(bestOverlapGoodness, (double)y.goodness);
			if (d == 0) bestClusterOverlap = hi - lo;
		}
		uint32_t similar = 0;
		for (uint32_t d = 0; d < nclusters; d++)
			if (d != c && (double)cl[d].goodness >= 0.9 * (double)x.goodness && (double)cl[d].goodness * 0.9 <= (double)x.goodness) similar++;
		const double mean = x.diagSum / n;
		const double var = std::max(0.0, x.diagSq / n - mean * mean);
		const uint64_t mid = (x.spanFrom + x.spanTo) / 2;
		const uint64_t wFrom = mid > 256 ? mid - 256 : 0, wTo = std::min<uint64_t>(len, mid + 256);
		const uint64_t sFrom = x.firstPos > 32 ? x.firstPos - 32 : 0, sTo = std::min<uint64_t>(len, x.firstPos + 32);
		f[0] = lg(len) / 16;
		f[1] = lg((double)seeds.size()) / 12;
		f[2] = lg(nclusters) / 10;
		f[3] = (float)(best / L);
		f[4] = lg(n) / 8;
		f[5] = (float)(x.matchBp / L);
		f[6] = lg((double)x.matchBp) / 16;
		f[7] = best ? (float)((double)x.goodness / best) : 0.0f;
		f[8] = (float)c / nclusters;
		f[9] = lg(c) / 10;
		f[10] = lg((double)(x.maxDiag - x.minDiag)) / 12;
		f[11] = lg(std::sqrt(var)) / 10;
		f[12] = (float)(span / L);
		f[13] = (float)(x.minPos / L);
		f[14] = (float)(x.maxPos / L);
		f[15] = (float)(std::fabs((double)mid - L / 2) / L);
		f[16] = maxCount ? (float)(x.rawSum / n / maxCount) : 0.0f;
		f[17] = maxCount ? (float)((double)x.minRaw / maxCount) : 0.0f;
		f[18] = x.reverse ? 1.0f : 0.0f;
		f[19] = (float)(x.inSum / n / 4);
		f[20] = (float)x.inMax / 4;
		f[21] = (float)(x.outSum / n / 4);
		f[22] = (float)x.outMax / 4;
		f[23] = (float)(x.inMulti / n);
		f[24] = (float)(x.outMulti / n);
		f[25] = lg(x.sccMax) / 20;
		f[26] = (float)(x.inScc / n);
		f[27] = lg((double)x.chainLen) / 24;
		f[28] = (float)(kmerEntropy(seq, wFrom, wTo, 2) / 4);
		f[29] = (float)(kmerEntropy(seq, sFrom, sTo, 3) / 6);
		f[30] = std::min(1.0f, (float)longestRun(seq, wFrom, wTo) / 16);
		f[31] = span > 0 ? (float)(spanCovered / span) : 0.0f;
		f[32] = lg(overlapping) / 8;
		f[33] = span > 0 ? (float)(bestClusterOverlap / span) : 0.0f;
		f[34] = bestOverlapGoodness > 0 ? (float)std::min(2.0, (double)x.goodness / bestOverlapGoodness) : 1.0f;
		f[35] = span > 0 ? lg(n / span * 100) / 8 : 0.0f;
		f[36] = maxSize ? (float)(n / maxSize) : 0.0f;
		f[37] = c == 0 ? 1.0f : 0.0f;
		f[38] = (double)x.goodness >= 0.9 * best ? 1.0f : 0.0f;
		f[39] = (float)(std::min<double>(x.minPos, L - x.maxPos) / L);
		f[40] = lg((double)x.minPos) / 16;
		f[41] = lg(L - (double)x.maxPos) / 16;
		f[42] = span > 0 ? (float)std::min(4.0, (double)(x.maxDiag - x.minDiag) / span) : 0.0f;
		f[43] = (float)(coveredCount / L);
		f[44] = (float)similar / nclusters;
		f[45] = seeds.empty() ? 0.0f : (float)(n / seeds.size());
		f[46] = lg((double)maxCount) / 16;
		f[47] = 1.0f;
		{
			// add [spanFrom, spanTo) to the union, merging the intervals it touches
			const uint64_t from = std::min<uint64_t>(x.spanFrom, len), to = std::min<uint64_t>(x.spanTo, len);
			if (to > from)
			{
				coveredCount += (to - from) - coveredIn(from, to);
				std::vector<std::pair<uint64_t, uint64_t>> merged;
				uint64_t lo = from, hi = to;
				for (const auto& iv : coveredBefore)
				{
					if (iv.second < lo || iv.first > hi)
						merged.push_back(iv);
					else
					{
						lo = std::min(lo, iv.first);
						hi = std::max(hi, iv.second);
					}
				}
				merged.emplace_back(lo, hi);
				std::sort(merged.begin(), merged.end());
				coveredBefore.swap(merged);
			}
		}
	}
}

void applyProposal(const SeedProposal& p, const GraphHost& g, uint64_t maxCount, size_t nreads, std::vector<SeedHit>* const* seeds, const char* const* seqs, const uint32_t* lens, int threads, ProposalStats& stats)
{
	if (!p.scorer || nreads == 0) return;
	auto t0 = std::chrono::steady_clock::now();
	std::vector<std::vector<uint32_t>> clusterOf(nreads);
	std::vector<uint32_t> ncl(nreads, 0);
	std::vector<uint64_t> rowStart(nreads + 1, 0);
	forEachRead(nreads, threads, [&](size_t i) {
		if (seeds[i] && !seeds[i]->empty()) ncl[i] = seedClusters(g, *seeds[i], clusterOf[i]);
	});
	// clusters scored per read: the first maxClusters (0: all)
	std::vector<uint32_t> scored(nreads);
	for (size_t i = 0; i < nreads; i++) scored[i] = p.maxClusters ? std::min(ncl[i], p.maxClusters) : ncl[i];
	for (size_t i = 0; i < nreads; i++) rowStart[i + 1] = rowStart[i] + scored[i];
	const uint64_t rows = rowStart[nreads];
	if (rows == 0) return;
	std::vector<float> feats(rows * kClusterFeatures);
	forEachRead(nreads, threads, [&](size_t i) {
		if (ncl[i]) clusterFeatures(g, maxCount, seqs[i], lens[i], *seeds[i], clusterOf[i], ncl[i], scored[i], feats.data() + rowStart[i] * kClusterFeatures);
	});
	std::vector<float> prob(rows * kProposalOutputs);
	p.scorer->score(feats.data(), rows, prob.data());
	std::vector<uint64_t> prunedSeeds(nreads, 0), prunedClusters(nreads, 0);
	forEachRead(nreads, threads, [&](size_t i) {
		if (!ncl[i]) return;
		std::vector<uint8_t> keep(ncl[i]);
		for (uint32_t c = 0; c < ncl[i]; c++)
		{
			keep[c] = c < p.keepTop || (c < scored[i] && prob[(rowStart[i] + c) * kProposalOutputs] >= p.threshold);
			prunedClusters[i] += !keep[c];
		}
		std::vector<SeedHit>& s = *seeds[i];
		size_t w = 0;
		for (size_t k = 0; k < s.size(); k++)
			if (keep[clusterOf[i][k]]) s[w++] = s[k];
		prunedSeeds[i] = s.size() - w;
		s.resize(w);
	});
	stats.clusters += rows;
	for (size_t i = 0; i < nreads; i++)
	{
		stats.prunedSeeds += prunedSeeds[i];
		stats.prunedClusters += prunedClusters[i];
	}
	stats.ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void ClusterRecords::add(int64_t read, const std::vector<float>& feats, uint32_t nclusters, const std::vector<uint32_t>& clusterOf, const std::vector<Alignment>& all, const std::vector<Alignment>& kept)
{
	std::vector<float> lab((size_t)nclusters * kProposalOutputs, 0.0f);
	for (const Alignment& a : all)
		if (a.seedIndex < clusterOf.size()) lab[(size_t)clusterOf[a.seedIndex] * kProposalOutputs + 1] = 1.0f;
	for (const Alignment& a : kept)
		if (a.seedIndex < clusterOf.size()) lab[(size_t)clusterOf[a.seedIndex] * kProposalOutputs + 0] = 1.0f;
	std::lock_guard<std::mutex> lock(m);
	features.insert(features.end(), feats.begin(), feats.begin() + (size_t)nclusters * kClusterFeatures);
	labels.insert(labels.end(), lab.begin(), lab.end());
	for (uint32_t c = 0; c < nclusters; c++)
	{
		reads.push_back(read);
		clusters.push_back((int32_t)c);
	}
}

void ClusterRecords::sort()
{
	std::lock_guard<std::mutex> lock(m);
	const size_t n = reads.size();
	std::vector<size_t> order(n);
	for (size_t i = 0; i < n; i++) order[i] = i;
	std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return reads[a] != reads[b] ? reads[a] < reads[b] : clusters[a] < clusters[b]; });
	std::vector<float> f(features.size()), l(labels.size());
	std::vector<int64_t> r(n);
	std::vector<int32_t> c(n);
	for (size_t k = 0; k < n; k++)
	{
		const size_t i = order[k];
		std::copy(features.begin() + i * kClusterFeatures, features.begin() + (i + 1) * kClusterFeatures, f.begin() + k * kClusterFeatures);
		std::copy(labels.begin() + i * kProposalOutputs, labels.begin() + (i + 1) * kProposalOutputs, l.begin() + k * kProposalOutputs);
		r[k] = reads[i];
		c[k] = clusters[i];
	}
	features.swap(f);
	labels.swap(l);
	reads.swap(r);
	clusters.swap(c);
}

} // namespace ga
