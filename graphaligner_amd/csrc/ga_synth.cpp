// Synthetic workload generator: seeded variation graphs in GFA and reads sampled
// from random graph paths. Deterministic for a given seed (std::mt19937_64 with
// hand-written distributions so the output does not depend on the C++ library).
//
// Variation graph: a uniform ACGT genome with SNP bubbles, short indel bubbles and
// optional structural-variant bubbles (insertion / deletion / inversion). Backbone
// sequence between variant sites is chopped into nodes of at most `nodeMax` bp and
// every edge is `L a <o> b <o> 0M`.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

struct Rng
{
	std::mt19937_64 gen;
	explicit Rng(uint64_t seed) : gen(seed) {}
	uint64_t next() { return gen(); }
	// uniform integer in [lo, hi]
	int64_t uniform(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }
	double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
	double normal()
	{
		double u1 = unit();
		double u2 = unit();
		if (u1 < 1e-300) u1 = 1e-300;
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
	}
	char base() { return "ACGT"[next() & 3]; }
};

char complement(char c)
{
	switch (c)
	{
		case 'A': return 'T';
		case 'C': return 'G';
		case 'G': return 'C';
		case 'T': return 'A';
	}
	return 'N';
}

std::string revcomp(const std::string& s)
{
	std::string r(s.rbegin(), s.rend());
	for (auto& c : r) c = complement(c);
	return r;
}

enum SiteType { SNP = 0, INS = 1, DEL = 2, INV = 3 };

struct Site
{
	int64_t pos;     // first affected genome position
	int64_t len;     // reference bases spanned (0 for insertions)
	int type;
	std::string alt; // alternative allele (SNP base, inserted sequence)
};

struct End
{
	int64_t node;
	bool plus;
};

struct Builder
{
	int nodeMax;
	std::vector<std::string> nodes;
	std::vector<std::string> edges;
	std::pair<int64_t, int64_t> chain(const std::string& seq)
	{
		int64_t first = -1, last = -1;
		for (size_t off = 0; off < seq.size(); off += nodeMax)
		{
			int64_t id = nodes.size() + 1;
			nodes.push_back(seq.substr(off, nodeMax));
			if (last != -1) edge(End{last, true}, End{id, true});
			if (first == -1) first = id;
			last = id;
		}
		return {first, last};
	}
	void edge(End from, End to)
	{
		char buf[96];
		snprintf(buf, sizeof(buf), "L\t%lld\t%c\t%lld\t%c\t0M", (long long)from.node, from.plus ? '+' : '-', (long long)to.node, to.plus ? '+' : '-');
		edges.push_back(buf);
	}
	void connect(const std::vector<End>& tails, End head)
	{
		for (auto t : tails) edge(t, head);
	}
};

struct Workload
{
	std::string genome;
	std::vector<Site> sites;
};

// ---- compacted de Bruijn graph -------------------------------------------

// k-mers of up to 63 bases, two bits per base (A0 C1 G2 T3), first base in the
// highest bits
typedef unsigned __int128 Kmer;

int baseCode(char c)
{
	switch (c)
	{
		case 'A': return 0;
		case 'C': return 1;
		case 'G': return 2;
		case 'T': return 3;
	}
	return -1;
}

// An oriented k-mer with its reverse complement, so that stepping to a neighbour
// keeps both in O(1).
struct Oriented
{
	Kmer fw, rc;
};

struct Dbg
{
	int k;
	Kmer mask;
	std::vector<Kmer> canon; // sorted canonical k-mers
	int64_t find(Kmer x) const
	{
		auto it = std::lower_bound(canon.begin(), canon.end(), x);
		return it != canon.end() && *it == x ? (int64_t)(it - canon.begin()) : -1;
	}
	// canonical index and whether the oriented k-mer reads as the canonical one
	int64_t id(const Oriented& o, bool& forward) const
	{
		forward = o.fw <= o.rc;
		return find(forward ? o.fw : o.rc);
	}
	Oriented step(const Oriented& o, int c) const
	{
		Oriented n;
		n.fw = ((o.fw << 2) | (Kmer)c) & mask;
		n.rc = (o.rc >> 2) | ((Kmer)(3 - c) << (2 * (k - 1)));
		return n;
	}
	static Oriented flip(const Oriented& o) { return Oriented{o.rc, o.fw}; }
	// the existing k-mers following o; returns their number (out holds up to 4)
	int successors(const Oriented& o, Oriented* out) const
	{
		int n = 0;
		for (int c = 0; c < 4; c++)
		{
			Oriented x = step(o, c);
			bool f;
			if (id(x, f) >= 0) out[n++] = x;
		}
		return n;
	}
	int predecessors(const Oriented& o, Oriented* out) const
	{
		Oriented tmp[4];
		int n = successors(flip(o), tmp);
		for (int i = 0; i < n; i++) out[i] = flip(tmp[i]);
		return n;
	}
	Oriented of(int64_t i) const
	{
		Oriented o;
		o.fw = canon[i];
		o.rc = 0;
		Kmer x = canon[i];
		for (int b = 0; b < k; b++)
		{
			o.rc = (o.rc << 2) | (Kmer)(3 - (int)(x & 3));
			x >>= 2;
		}
		return o;
	}
	std::string text(Kmer x) const
	{
		std::string s(k, 'A');
		for (int b = k - 1; b >= 0; b--)
		{
			s[b] = "ACGT"[(int)(x & 3)];
			x >>= 2;
		}
		return s;
	}
};

} // namespace

extern "C" {

struct GaSynthGraphConfig
{
	int64_t genomeLength;
	int64_t snpMin, snpMax;       // SNP spacing U[snpMin, snpMax]
	int64_t indelMin, indelMax;   // indel spacing
	int64_t indelLenMin, indelLenMax;
	int64_t svEvery;              // 0 = no structural variants
	int64_t svLenMin, svLenMax;   // log-uniform SV length
	int32_t nodeMax;
	uint64_t seed;
};

struct GaSynthReadConfig
{
	int64_t numReads;
	int32_t lengthModel;          // 0 uniform [lenA, lenB]; 1 log-normal (median lenA, sigma) clipped [lenMin, lenMax]
	int64_t lenA, lenB;
	double sigma;
	int64_t lenMin, lenMax;
	double errorMin, errorMax;    // per-read error rate U[errorMin, errorMax]
	uint64_t seed;
};

static Workload* makeWorkload(const GaSynthGraphConfig* cfg)
{
	Rng rng(cfg->seed);
	auto* w = new Workload;
	w->genome.resize(cfg->genomeLength);
	for (auto& c : w->genome) c = rng.base();
	std::vector<Site> all;
	for (int64_t p = rng.uniform(cfg->snpMin, cfg->snpMax); p < cfg->genomeLength - 1; p += rng.uniform(cfg->snpMin, cfg->snpMax))
	{
		char alt;
		do alt = rng.base(); while (alt == w->genome[p]);
		all.push_back(Site{p, 1, SNP, std::string(1, alt)});
	}
	for (int64_t p = rng.uniform(cfg->indelMin, cfg->indelMax); p < cfg->genomeLength - 20; p += rng.uniform(cfg->indelMin, cfg->indelMax))
	{
		int64_t len = rng.uniform(cfg->indelLenMin, cfg->indelLenMax);
		if (rng.next() & 1)
		{
			std::string ins;
			for (int64_t i = 0; i < len; i++) ins += rng.base();
			all.push_back(Site{p, 0, INS, ins});
		}
		else
		{
			all.push_back(Site{p, len, DEL, ""});
		}
	}
	if (cfg->svEvery > 0)
	{
		for (int64_t p = cfg->svEvery / 2 + rng.uniform(0, cfg->svEvery / 4); p < cfg->genomeLength - cfg->svLenMax - 100; p += cfg->svEvery + rng.uniform(-cfg->svEvery / 4, cfg->svEvery / 4))
		{
			double lo = std::log((double)cfg->svLenMin), hi = std::log((double)cfg->svLenMax);
			int64_t len = (int64_t)std::exp(lo + (hi - lo) * rng.unit());
			int t = rng.uniform(0, 2);
			if (t == 0)
			{
				std::string ins;
				for (int64_t i = 0; i < len; i++) ins += rng.base();
				all.push_back(Site{p, 0, INS, ins});
			}
			else if (t == 1) all.push_back(Site{p, len, DEL, ""});
			else all.push_back(Site{p, len, INV, ""});
		}
	}
	std::stable_sort(all.begin(), all.end(), [](const Site& a, const Site& b) { return a.pos < b.pos; });
	// keep sites separated by at least two backbone bases
	int64_t lastEnd = 0;
	for (auto& s : all)
	{
		if (s.pos < lastEnd + 2) continue;
		if (s.pos + s.len > cfg->genomeLength - 2) continue;
		w->sites.push_back(s);
		lastEnd = s.pos + s.len;
	}
	return w;
}

static void buildGraph(const Workload& w, int nodeMax, Builder& b)
{
	b.nodeMax = nodeMax;
	std::vector<End> tails;
	int64_t pos = 0;
	auto backbone = [&](int64_t to) {
		if (to <= pos) return;
		auto c = b.chain(w.genome.substr(pos, to - pos));
		b.connect(tails, End{c.first, true});
		tails = {End{c.second, true}};
		pos = to;
	};
	for (const auto& s : w.sites)
	{
		backbone(s.pos);
		if (s.type == SNP)
		{
			auto r = b.chain(w.genome.substr(s.pos, 1));
			auto a = b.chain(s.alt);
			b.connect(tails, End{r.first, true});
			b.connect(tails, End{a.first, true});
			tails = {End{r.second, true}, End{a.second, true}};
			pos = s.pos + 1;
		}
		else if (s.type == INS)
		{
			auto a = b.chain(s.alt);
			b.connect(tails, End{a.first, true});
			tails.push_back(End{a.second, true});
		}
		else if (s.type == DEL)
		{
			auto r = b.chain(w.genome.substr(s.pos, s.len));
			auto old = tails;
			b.connect(tails, End{r.first, true});
			tails = {End{r.second, true}};
			for (auto t : old) tails.push_back(t);
			pos = s.pos + s.len;
		}
		else
		{
			auto r = b.chain(w.genome.substr(s.pos, s.len));
			b.connect(tails, End{r.first, true});
			b.connect(tails, End{r.second, false});
			tails = {End{r.second, true}, End{r.first, false}};
			pos = s.pos + s.len;
		}
	}
	backbone(w.genome.size());
}

// Haplotype sequence starting at genome position `start`, choosing an allele per
// site at random, until `need` bases are produced or the genome ends.
static std::string haplotype(const Workload& w, int64_t start, int64_t need, Rng& rng)
{
	std::string out;
	auto it = std::lower_bound(w.sites.begin(), w.sites.end(), start, [](const Site& s, int64_t p) { return s.pos < p; });
	// a start inside a site's span moves to the site's end
	if (it != w.sites.begin())
	{
		auto prev = it - 1;
		if (prev->pos + std::max<int64_t>(prev->len, 1) > start) start = prev->pos + std::max<int64_t>(prev->len, 1);
	}
	int64_t pos = start;
	while ((int64_t)out.size() < need && pos < (int64_t)w.genome.size())
	{
		int64_t to = it == w.sites.end() ? (int64_t)w.genome.size() : it->pos;
		if (to > pos) out += w.genome.substr(pos, to - pos);
		pos = std::max(pos, to);
		if (it == w.sites.end()) break;
		bool alt = rng.next() & 1;
		const Site& s = *it;
		if (s.type == SNP) out += alt ? s.alt : w.genome.substr(s.pos, 1);
		else if (s.type == INS) { if (alt) out += s.alt; }
		else if (s.type == DEL) { if (!alt) out += w.genome.substr(s.pos, s.len); }
		else out += alt ? revcomp(w.genome.substr(s.pos, s.len)) : w.genome.substr(s.pos, s.len);
		pos = s.pos + s.len;
		++it;
	}
	if ((int64_t)out.size() > need) out.resize(need);
	return out;
}

static std::string mutate(const std::string& s, double rate, Rng& rng)
{
	std::string out;
	out.reserve(s.size() + s.size() / 8);
	for (char c : s)
	{
		if (rng.unit() >= rate)
		{
			out += c;
			continue;
		}
		int64_t kind = rng.uniform(0, 9); // substitution : insertion : deletion = 4 : 3 : 3
		if (kind < 4)
		{
			char b;
			do b = rng.base(); while (b == c);
			out += b;
		}
		else if (kind < 7)
		{
			out += rng.base();
			out += c;
		}
	}
	return out;
}

// Reads sampled from the strings of `sources` (one picked per read with
// probability proportional to its length), written to <prefix>.fa.
static bool writeReads(const std::vector<const std::string*>& sources, const GaSynthReadConfig* rcfg, const std::string& path)
{
	FILE* r = fopen(path.c_str(), "w");
	if (!r) return false;
	Rng rng(rcfg->seed);
	int64_t total = 0;
	for (auto* s : sources) total += s->size();
	for (int64_t i = 0; i < rcfg->numReads; i++)
	{
		int64_t len;
		if (rcfg->lengthModel == 0) len = rng.uniform(rcfg->lenA, rcfg->lenB);
		else
		{
			double v = std::exp(std::log((double)rcfg->lenA) + rcfg->sigma * rng.normal());
			len = std::max(rcfg->lenMin, std::min(rcfg->lenMax, (int64_t)v));
		}
		int64_t at = rng.uniform(0, total - 1);
		size_t src = 0;
		while (at >= (int64_t)sources[src]->size())
		{
			at -= sources[src]->size();
			src++;
		}
		const std::string& g = *sources[src];
		len = std::min<int64_t>(len, g.size());
		int64_t start = rng.uniform(0, (int64_t)g.size() - len);
		std::string piece = g.substr(start, len);
		if (rng.next() & 1) piece = revcomp(piece);
		double rate = rcfg->errorMin + (rcfg->errorMax - rcfg->errorMin) * rng.unit();
		fprintf(r, ">read%lld\n%s\n", (long long)i, mutate(piece, rate, rng).c_str());
	}
	fclose(r);
	return true;
}

// A de Bruijn graph workload: a seeded uniform genome with inserted copies of
// repeat families (each copy substituted at its own rate, either strand) and
// tandem arrays, compacted into the unitigs of its canonical k-mers.
struct GaSynthDbgConfig
{
	int64_t genomeLength;
	int64_t repeatCopies;              // copies inserted in all
	int64_t familyLenMin, familyLenMax;
	int64_t copiesMin, copiesMax;      // copies per family
	double divergenceMax;              // per-copy substitution rate U[0, max]
	int64_t tandemArrays;
	int64_t periodMin, periodMax;      // tandem period
	int64_t unitsMin, unitsMax;        // tandem units per array
	double tandemDivergence;           // per-unit substitution rate
	int32_t k;                         // <= 63
	uint64_t seed;
};

// Writes <prefix>.gfa (unitigs with k-1 overlaps) and <prefix>.fa. Returns the
// number of unitigs, or -1.
int64_t ga_synth_dbg_workload(const GaSynthDbgConfig* cfg, const GaSynthReadConfig* rcfg, const char* prefix)
{
	if (cfg->k < 3 || cfg->k > 63 || cfg->k % 2 == 0) return -1;
	Rng rng(cfg->seed);
	std::string genome(cfg->genomeLength, 'A');
	for (auto& c : genome) c = rng.base();
	auto substitute = [&](std::string x, double rate) {
		for (auto& c : x)
			if (rng.unit() < rate)
			{
				char b;
				do b = rng.base(); while (b == c);
				c = b;
			}
		return x;
	};
	// insertions are collected as (position, sequence) and applied from the
	// back so that earlier positions stay valid
	std::vector<std::pair<int64_t, std::string>> inserts;
	for (int64_t placed = 0; placed < cfg->repeatCopies;)
	{
		std::string family(rng.uniform(cfg->familyLenMin, cfg->familyLenMax), 'A');
		for (auto& c : family) c = rng.base();
		const int64_t copies = std::min(rng.uniform(cfg->copiesMin, cfg->copiesMax), cfg->repeatCopies - placed);
		for (int64_t i = 0; i < copies; i++)
		{
			std::string copy = substitute(family, cfg->divergenceMax * rng.unit());
			if (rng.next() & 1) copy = revcomp(copy);
			inserts.emplace_back(rng.uniform(0, cfg->genomeLength), copy);
		}
		placed += copies;
	}
	for (int64_t a = 0; a < cfg->tandemArrays; a++)
	{
		std::string unit(rng.uniform(cfg->periodMin, cfg->periodMax), 'A');
		for (auto& c : unit) c = rng.base();
		std::string array;
		const int64_t units = rng.uniform(cfg->unitsMin, cfg->unitsMax);
		for (int64_t i = 0; i < units; i++) array += substitute(unit, cfg->tandemDivergence);
		inserts.emplace_back(rng.uniform(0, cfg->genomeLength), array);
	}
	std::stable_sort(inserts.begin(), inserts.end(), [](const std::pair<int64_t, std::string>& a, const std::pair<int64_t, std::string>& b) { return a.first > b.first; });
	for (const auto& in : inserts) genome.insert((size_t)in.first, in.second);

	Dbg d;
	d.k = cfg->k;
	d.mask = (((Kmer)1) << (2 * d.k)) - 1;
	{
		Kmer fw = 0, rc = 0;
		d.canon.reserve(genome.size());
		for (size_t i = 0; i < genome.size(); i++)
		{
			const int c = baseCode(genome[i]);
			fw = ((fw << 2) | (Kmer)c) & d.mask;
			rc = (rc >> 2) | ((Kmer)(3 - c) << (2 * (d.k - 1)));
			if (i + 1 >= (size_t)d.k) d.canon.push_back(fw <= rc ? fw : rc);
		}
		std::sort(d.canon.begin(), d.canon.end());
		d.canon.erase(std::unique(d.canon.begin(), d.canon.end()), d.canon.end());
	}
	const int64_t N = d.canon.size();
	// maximal non-branching paths: a k-mer joins its successor when it has one
	// successor and that successor has one predecessor
	std::vector<int32_t> unitigOf(N, -1);
	struct Unitig
	{
		Oriented first, last;
		std::string seq;
	};
	std::vector<Unitig> unitigs;
	Oriented nb[4];
	for (int64_t i = 0; i < N; i++)
	{
		if (unitigOf[i] >= 0) continue;
		Oriented cur = d.of(i);
		// back to the path's start (stopping on a cycle back to i)
		while (true)
		{
			if (d.predecessors(cur, nb) != 1) break;
			Oriented pr = nb[0];
			Oriented tmp[4];
			if (d.successors(pr, tmp) != 1) break;
			bool f;
			if (d.id(pr, f) == i) break;
			cur = pr;
		}
		Unitig u;
		u.first = cur;
		u.seq = d.text(cur.fw);
		bool f;
		const int32_t uid = (int32_t)unitigs.size();
		unitigOf[d.id(cur, f)] = uid;
		while (true)
		{
			if (d.successors(cur, nb) != 1) break;
			Oriented nx = nb[0];
			Oriented tmp[4];
			if (d.predecessors(nx, tmp) != 1) break;
			const int64_t nid = d.id(nx, f);
			if (unitigOf[nid] >= 0) break;
			unitigOf[nid] = uid;
			u.seq += "ACGT"[(int)(nx.fw & 3)];
			cur = nx;
		}
		u.last = cur;
		unitigs.push_back(std::move(u));
	}
	// edges between unitig ends; each edge and its reverse written once
	std::string gfaPath = std::string(prefix) + ".gfa";
	FILE* f = fopen(gfaPath.c_str(), "w");
	if (!f) return -1;
	fprintf(f, "H\tVN:Z:1.0\n");
	for (size_t u = 0; u < unitigs.size(); u++) fprintf(f, "S\t%zu\t%s\n", u + 1, unitigs[u].seq.c_str());
	auto sideOf = [&](const Oriented& x, int32_t& unit, bool& plus) {
		bool fwd;
		const int64_t id = d.id(x, fwd);
		unit = unitigOf[id];
		const Unitig& t = unitigs[unit];
		// x starts unit on its + strand, or is the reverse of its last k-mer
		plus = x.fw == t.first.fw;
	};
	for (size_t u = 0; u < unitigs.size(); u++)
	{
		for (int side = 0; side < 2; side++)
		{
			// side 0: leaving u on +, after its last k-mer; side 1: leaving u on -
			const Oriented from = side == 0 ? unitigs[u].last : Dbg::flip(unitigs[u].first);
			const int n = d.successors(from, nb);
			for (int e = 0; e < n; e++)
			{
				int32_t v;
				bool vplus;
				sideOf(nb[e], v, vplus);
				const bool uplus = side == 0;
				// the same edge read backwards is (v, !vplus) -> (u, !uplus)
				const std::pair<int64_t, int> a((int64_t)u, uplus ? 0 : 1), b((int64_t)v, vplus ? 1 : 0);
				if (b < a) continue;
				fprintf(f, "L\t%zu\t%c\t%d\t%c\t%dM\n", u + 1, uplus ? '+' : '-', v + 1, vplus ? '+' : '-', d.k - 1);
			}
		}
	}
	fclose(f);
	if (rcfg && rcfg->numReads > 0 && !writeReads({&genome}, rcfg, std::string(prefix) + ".fa")) return -1;
	return (int64_t)unitigs.size();
}

// Writes <prefix>.gfa and <prefix>.fa. Returns the number of graph nodes, or -1.
int64_t ga_synth_variation_workload(const GaSynthGraphConfig* gcfg, const GaSynthReadConfig* rcfg, const char* prefix)
{
	Workload* w = makeWorkload(gcfg);
	Builder b;
	buildGraph(*w, gcfg->nodeMax, b);
	std::string gfaPath = std::string(prefix) + ".gfa";
	FILE* f = fopen(gfaPath.c_str(), "w");
	if (!f)
	{
		delete w;
		return -1;
	}
	fprintf(f, "H\tVN:Z:1.0\n");
	for (size_t i = 0; i < b.nodes.size(); i++) fprintf(f, "S\t%zu\t%s\n", i + 1, b.nodes[i].c_str());
	for (const auto& e : b.edges) fprintf(f, "%s\n", e.c_str());
	fclose(f);
	if (rcfg && rcfg->numReads > 0)
	{
		std::string faPath = std::string(prefix) + ".fa";
		FILE* r = fopen(faPath.c_str(), "w");
		if (!r)
		{
			delete w;
			return -1;
		}
		Rng rng(rcfg->seed);
		int64_t G = w->genome.size();
		for (int64_t i = 0; i < rcfg->numReads; i++)
		{
			int64_t len;
			if (rcfg->lengthModel == 0) len = rng.uniform(rcfg->lenA, rcfg->lenB);
			else
			{
				double v = std::exp(std::log((double)rcfg->lenA) + rcfg->sigma * rng.normal());
				len = (int64_t)v;
				len = std::max(rcfg->lenMin, std::min(rcfg->lenMax, len));
			}
			len = std::min<int64_t>(len, G / 2);
			std::string hap;
			for (int attempt = 0; attempt < 100 && (int64_t)hap.size() < len; attempt++)
			{
				int64_t start = rng.uniform(0, std::max<int64_t>(0, G - len - len / 4 - 1));
				hap = haplotype(*w, start, len, rng);
			}
			if (rng.next() & 1) hap = revcomp(hap);
			double rate = rcfg->errorMin + (rcfg->errorMax - rcfg->errorMin) * rng.unit();
			std::string read = mutate(hap, rate, rng);
			fprintf(r, ">read%lld\n%s\n", (long long)i, read.c_str());
		}
		fclose(r);
	}
	int64_t n = b.nodes.size();
	delete w;
	return n;
}

} // extern "C"
