// Host-side alignment graph construction (see ga_graph.hpp).
#include "ga_graph.hpp"

#include <algorithm>
#include <atomic>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <tuple>
#include <unordered_set>

namespace ga {
namespace {

constexpr uint32_t kSplit = 64;
constexpr size_t kNone = (size_t)-1;

struct Side
{
	int id = 0;
	bool plus = false;
	bool operator==(const Side& o) const { return id == o.id && plus == o.plus; }
};
struct SideHash
{
	size_t operator()(const Side& s) const { return std::hash<int>()(s.id) ^ std::hash<bool>()(s.plus); }
};
struct SidePairHash
{
	// Cantor pairing evaluated in double, as src/GfaGraph.h:34-44 does
	size_t operator()(const std::pair<Side, Side>& x) const
	{
		size_t pairing = .5 * (x.first.id + x.second.id) * (x.first.id + x.second.id + 1) + x.second.id;
		return std::hash<size_t>()(pairing);
	}
};

// GFA S/L records in the container layout that decides node numbering
// (src/GfaGraph.h:64-71).
struct Gfa
{
	std::unordered_map<int, std::string> nodes;
	std::unordered_map<Side, std::vector<Side>, SideHash> edges;
	std::unordered_map<std::pair<Side, Side>, size_t, SidePairHash> overlaps;
	size_t edgeOverlap = (size_t)-1;
	std::unordered_map<int, std::string> names;
};

int nameId(std::unordered_map<std::string, int>& ids, const std::string& name)
{
	auto it = ids.find(name);
	if (it != ids.end()) return it->second;
	int id = ids.size();
	ids[name] = id;
	return id;
}

// src/GfaGraph.cpp:218-396 with varying overlaps allowed (src/Aligner.cpp:640)
Gfa parseGfa(std::istream& in)
{
	std::unordered_map<std::string, int> ids;
	Gfa g;
	bool varying = false;
	while (in.good())
	{
		std::string line;
		std::getline(in, line);
		if (!in.good()) break; // a last line without newline is not read
		if (line.empty() || (line[0] != 'S' && line[0] != 'L')) continue;
		std::stringstream ss{line};
		std::string tag;
		ss >> tag;
		if (line[0] == 'S')
		{
			std::string name, seq;
			ss >> name;
			int id = nameId(ids, name);
			ss >> seq;
			if (seq == "*") throw std::runtime_error("GFA nodes without sequence (*) are not supported");
			g.nodes[id] = seq;
		}
		else
		{
			std::string fromName, toName, fromOri, toOri;
			ss >> fromName;
			int from = nameId(ids, fromName);
			ss >> fromOri >> toName;
			int to = nameId(ids, toName);
			ss >> toOri;
			ss.get();
			int overlap = 0;
			if (ss.peek() != '*')
			{
				ss >> overlap;
				char unit = 'z';
				ss >> unit;
			}
			if (overlap < 0) throw std::runtime_error("GFA edge overlap cannot be negative");
			if (g.edgeOverlap != (size_t)-1 && (size_t)overlap != g.edgeOverlap) varying = true;
			g.edgeOverlap = overlap;
			Side a{from, fromOri == "+"}, b{to, toOri == "+"};
			g.edges[a].push_back(b);
			g.overlaps[std::make_pair(a, b)] = overlap;
		}
	}
	if (varying || g.edges.empty()) g.edgeOverlap = 0;
	bool allInts = true;
	for (const auto& p : ids)
	{
		g.names[p.second] = p.first;
		if (!allInts) continue;
		char* end = nullptr;
		long v = strtol(p.first.c_str(), &end, 10);
		if (*end != '\0' || v <= INT32_MIN || v >= INT32_MAX) allInts = false;
	}
	for (const auto& p : g.overlaps)
	{
		auto a = g.nodes.find(p.first.first.id), b = g.nodes.find(p.first.second.id);
		if (a == g.nodes.end() || b == g.nodes.end()) continue;
		if (a->second.size() <= p.second || b->second.size() <= p.second) throw std::runtime_error("GFA edge overlap is larger than a node");
	}
	if (allInts)
	{
		// src/GfaGraph.cpp:182-216: renumber to the integer names
		std::unordered_map<int, std::string> nodes;
		std::unordered_map<Side, std::vector<Side>, SideHash> edges;
		std::unordered_map<std::pair<Side, Side>, size_t, SidePairHash> overlaps;
		auto num = [&](int id) { return std::stoi(g.names[id]); };
		for (auto p : g.overlaps)
		{
			auto k = p.first;
			k.first.id = num(k.first.id);
			k.second.id = num(k.second.id);
			overlaps[k] = p.second;
		}
		for (auto p : g.nodes) nodes[num(p.first)] = p.second;
		for (auto e : g.edges)
			for (auto t : e.second) edges[Side{num(e.first.id), e.first.plus}].push_back(Side{num(t.id), t.plus});
		g.overlaps = std::move(overlaps);
		g.nodes = std::move(nodes);
		g.edges = std::move(edges);
		g.names.clear();
	}
	std::vector<Side> dangling;
	for (auto& e : g.edges)
	{
		if (g.nodes.count(e.first.id) == 0)
		{
			dangling.push_back(e.first);
			continue;
		}
		for (size_t i = e.second.size() - 1; i < e.second.size() + 1; i--)
			if (g.nodes.count(e.second[i].id) == 0) e.second.erase(e.second.begin() + i);
	}
	for (auto d : dangling) g.edges.erase(g.edges.find(d));
	return g;
}

char complementBase(char c)
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
	throw std::runtime_error("invalid sequence character in graph");
}

// 4-bit IUPAC set of a graph base (bit0 A, bit1 C, bit2 G, bit3 T); 0 if invalid
uint32_t baseSet(char c)
{
	switch (c)
	{
		case 'a': case 'A': return 1;
		case 'c': case 'C': return 2;
		case 'g': case 'G': return 4;
		case 't': case 'T': case 'u': case 'U': return 8;
		case 'r': case 'R': return 1 | 4;
		case 'y': case 'Y': return 2 | 8;
		case 's': case 'S': return 4 | 2;
		case 'w': case 'W': return 1 | 8;
		case 'k': case 'K': return 4 | 8;
		case 'm': case 'M': return 1 | 2;
		case 'b': case 'B': return 2 | 4 | 8;
		case 'd': case 'D': return 1 | 4 | 8;
		case 'h': case 'H': return 1 | 2 | 8;
		case 'v': case 'V': return 1 | 2 | 4;
		case 'n': case 'N': return 15;
	}
	return 0;
}

// Graph under construction: adjacency as vectors (neighbour order matters).
struct Build
{
	std::vector<uint32_t> len;
	std::vector<int32_t> id;
	std::vector<uint32_t> off;
	std::vector<uint8_t> rev;
	std::vector<uint8_t> ambiguous;
	std::vector<std::array<uint64_t, 2>> seq;   // all nodes (ambiguous ones get zeros)
	std::vector<std::array<uint64_t, 4>> amb;   // all nodes
	std::vector<std::vector<size_t>> in, out;
	std::unordered_map<int, std::vector<size_t>> lookup;
	std::unordered_map<int, size_t> osize;
	std::unordered_map<int, std::string> oname;
	uint64_t bp = 0;

	void addPiece(int nodeId, size_t offset, const std::string& s, bool reverse)
	{
		bp += s.size();
		lookup[nodeId].push_back(len.size());
		len.push_back(s.size());
		id.push_back(nodeId);
		in.emplace_back();
		out.emplace_back();
		rev.push_back(reverse);
		off.push_back(offset);
		std::array<uint64_t, 2> packed{0, 0};
		std::array<uint64_t, 4> masks{0, 0, 0, 0};
		bool isAmb = false;
		for (size_t i = 0; i < s.size(); i++)
		{
			uint32_t set = baseSet(s[i]);
			if (set == 0) throw std::runtime_error("invalid sequence character in graph");
			for (int b = 0; b < 4; b++)
				if (set & (1u << b)) masks[b] |= (uint64_t)1 << i;
			int code = set == 1 ? 0 : set == 2 ? 1 : set == 4 ? 2 : set == 8 ? 3 : -1;
			if (code < 0) isAmb = true;
			else packed[i / 32] |= (uint64_t)code << ((i % 32) * 2);
		}
		ambiguous.push_back(isAmb);
		seq.push_back(isAmb ? std::array<uint64_t, 2>{0, 0} : packed);
		amb.push_back(masks);
	}

	// src/AlignmentGraph.cpp:56-97
	void addNode(int nodeId, const std::string& s, const std::string& name, bool reverse, const std::vector<size_t>& breaks)
	{
		if (lookup.count(nodeId) != 0) return;
		osize[nodeId] = s.size();
		oname[nodeId] = name;
		for (size_t b = 1; b < breaks.size(); b++)
		{
			if (breaks[b] == breaks[b - 1]) continue;
			for (size_t o = breaks[b - 1]; o < breaks[b]; o += kSplit)
			{
				size_t size = std::min<size_t>(kSplit, breaks[b] - o);
				addPiece(nodeId, o, s.substr(o, size), reverse);
				if (o > 0)
				{
					out[out.size() - 2].push_back(out.size() - 1);
					in[in.size() - 1].push_back(in.size() - 2);
				}
			}
		}
	}

	// src/AlignmentGraph.cpp:230-250
	void addEdge(int fromId, int toId, size_t startOffset)
	{
		size_t from = lookup.at(fromId).back();
		size_t to = kNone;
		for (auto n : lookup[toId])
			if (off[n] == startOffset) to = n;
		if (to == kNone) throw std::runtime_error("GFA edge points inside a node");
		if (std::find(in[to].begin(), in[to].end(), from) == in[to].end()) in[to].push_back(from);
		if (std::find(out[from].begin(), out[from].end(), to) == out[from].end()) out[from].push_back(to);
	}
};

size_t ufFind(std::vector<size_t>& parent, size_t x)
{
	size_t root = x;
	while (parent[root] != root) root = parent[root];
	while (parent[x] != root)
	{
		size_t next = parent[x];
		parent[x] = root;
		x = next;
	}
	return root;
}

void ufUnion(std::vector<size_t>& parent, std::vector<size_t>& rank, size_t a, size_t b)
{
	a = ufFind(parent, a);
	b = ufFind(parent, b);
	if (rank[a] < rank[b]) std::swap(a, b);
	parent[b] = a;
	if (rank[a] == rank[b]) rank[a] += 1;
}

struct Finalizer
{
	Build& g;
	size_t n;
	std::vector<uint8_t> lin;
	std::vector<size_t> comp, chain, chainPos;
	explicit Finalizer(Build& b) : g(b), n(b.len.size()) {}

	// src/AlignmentGraph.cpp:915-985
	size_t renumberAmbiguous()
	{
		std::vector<size_t> renum(n);
		size_t plain = 0, ambCount = 0;
		for (size_t i = 0; i < n; i++)
		{
			if (!g.ambiguous[i]) renum[i] = plain++;
			else renum[i] = n - 1 - ambCount++;
		}
		if (ambCount == 0) return plain;
		auto perm = [&](auto& v) {
			auto copy = v;
			for (size_t i = 0; i < n; i++) v[renum[i]] = copy[i];
		};
		perm(g.len);
		perm(g.off);
		perm(g.id);
		perm(g.in);
		perm(g.out);
		perm(g.rev);
		perm(g.seq);
		perm(g.amb);
		perm(g.ambiguous);
		for (auto& p : g.lookup)
			for (auto& x : p.second) x = renum[x];
		for (size_t i = 0; i < n; i++)
		{
			for (auto& x : g.in[i]) x = renum[x];
			for (auto& x : g.out[i]) x = renum[x];
		}
		return plain;
	}

	// src/AlignmentGraph.cpp:632-724 (see the oracle's note on the stale stack read)
	void linearizable()
	{
		lin.assign(n, 0);
		std::vector<uint8_t> checked(n, 0), onStack(n, 0);
		std::vector<size_t> stack;
		for (size_t node = 0; node < n; node++)
		{
			if (checked[node]) continue;
			checked[node] = 1;
			if (g.in[node].size() != 1) continue;
			stack.assign(1, node);
			onStack[node] = 1;
			while (true)
			{
				size_t top = stack.back();
				if (g.in[top].size() != 1 || checked[top])
				{
					for (size_t i = 0; i + 1 < stack.size(); i++)
					{
						checked[stack[i]] = 1;
						lin[stack[i]] = 1;
						onStack[stack[i]] = 0;
					}
					lin[top] = 0;
					checked[top] = 1;
					onStack[top] = 0;
					break;
				}
				size_t nb = g.in[top][0];
				if (nb == node)
				{
					for (auto s : stack)
					{
						checked[s] = 1;
						lin[s] = 0;
						onStack[s] = 0;
					}
					break;
				}
				if (onStack[nb])
				{
					size_t i = stack.size() - 1;
					for (; i > 0; i--)
					{
						if (stack[i] == nb) break;
						checked[stack[i]] = 1;
						lin[stack[i]] = 0;
						onStack[stack[i]] = 0;
					}
					for (size_t k = 0; k < i; k++)
					{
						checked[stack[k]] = 1;
						lin[stack[k]] = 1;
						onStack[stack[k]] = 0;
					}
					break;
				}
				stack.push_back(nb);
				onStack[nb] = 1;
			}
		}
	}

	// src/AlignmentGraph.cpp:987-1094: Tarjan with an explicit stack, components
	// numbered in reverse completion order
	void components()
	{
		const size_t unset = kNone;
		std::vector<size_t> index(n, unset), low(n, unset), stack;
		std::vector<uint8_t> onStack(n, 0);
		comp.assign(n, unset);
		struct Frame
		{
			size_t v;
			size_t next;
		};
		std::vector<Frame> calls;
		size_t counter = 0, nextComp = 0;
		for (size_t root = 0; root < n; root++)
		{
			if (index[root] != unset) continue;
			calls.push_back(Frame{root, 0});
			index[root] = low[root] = counter++;
			stack.push_back(root);
			onStack[root] = 1;
			while (!calls.empty())
			{
				Frame& f = calls.back();
				size_t v = f.v;
				bool descended = false;
				while (f.next < g.out[v].size())
				{
					size_t w = g.out[v][f.next];
					if (index[w] == unset)
					{
						index[w] = low[w] = counter++;
						stack.push_back(w);
						onStack[w] = 1;
						calls.push_back(Frame{w, 0});
						descended = true;
						break;
					}
					if (onStack[w]) low[v] = std::min(low[v], index[w]);
					f.next++;
				}
				if (descended) continue;
				if (low[v] == index[v])
				{
					size_t w;
					do
					{
						w = stack.back();
						stack.pop_back();
						onStack[w] = 0;
						comp[w] = nextComp;
					} while (w != v);
					nextComp++;
				}
				calls.pop_back();
				if (!calls.empty())
				{
					Frame& parent = calls.back();
					low[parent.v] = std::min(low[parent.v], low[v]);
					parent.next++;
				}
			}
		}
		for (auto& c : comp) c = nextComp - 1 - c;
	}

	std::pair<bool, size_t> findBubble(size_t start, const std::vector<uint8_t>& tip)
	{
		std::vector<size_t> S{start};
		std::unordered_set<size_t> visited, seen{start};
		while (!S.empty())
		{
			size_t v = S.back();
			S.pop_back();
			seen.erase(v);
			visited.insert(v);
			if (g.out[v].empty()) return {false, 0};
			for (size_t u : g.out[v])
			{
				if (tip[u] || u == v) continue;
				if (u == start) return {false, 0};
				seen.insert(u);
				bool pending = false;
				for (size_t w : g.in[u])
				{
					if (w == u) continue;
					if (!tip[w] && visited.count(w) == 0)
					{
						pending = true;
						break;
					}
				}
				if (!pending) S.push_back(u);
			}
			if (S.size() == 1 && seen.size() == 1 && seen.count(S[0]) == 1)
			{
				size_t t = S.back();
				for (size_t u : g.out[t])
					if (u == start) return {false, 0};
				return {true, t};
			}
		}
		return {false, 0};
	}

	// src/AlignmentGraph.cpp:298-630 (tip chaining, cycles, bubbles, approximate positions)
	void chains()
	{
		chain.resize(n);
		for (size_t i = 0; i < n; i++) chain[i] = i;
		std::vector<size_t> rank(n, 0);
		std::vector<uint8_t> tip(n, 0);
		for (const auto& p : g.lookup)
			for (size_t i = 1; i < p.second.size(); i++) ufUnion(chain, rank, p.second[0], p.second[i]);
		std::vector<size_t> order(n);
		for (size_t i = 0; i < n; i++) order[i] = i;
		std::sort(order.begin(), order.end(), [this](size_t a, size_t b) { return comp[a] < comp[b]; });
		size_t numComp = n ? comp[order.back()] + 1 : 0;
		std::vector<uint8_t> fwTip(numComp, 1), bwTip(numComp, 1);
		for (size_t k = n; k-- > 0;)
		{
			size_t i = order[k];
			if (!fwTip[comp[i]]) continue;
			for (auto nb : g.out[i])
				if (comp[nb] == comp[i] || !fwTip[comp[nb]])
				{
					fwTip[comp[i]] = 0;
					break;
				}
		}
		for (size_t k = n; k-- > 0;)
		{
			size_t i = order[k];
			if (fwTip[comp[i]])
				for (auto nb : g.out[i]) ufUnion(chain, rank, i, nb);
		}
		for (size_t k = 0; k < n; k++)
		{
			size_t i = order[k];
			if (!bwTip[comp[i]]) continue;
			for (auto nb : g.in[i])
				if (comp[nb] == comp[i] || !bwTip[comp[nb]])
				{
					bwTip[comp[i]] = 0;
					break;
				}
		}
		for (size_t k = 0; k < n; k++)
		{
			size_t i = order[k];
			if (bwTip[comp[i]])
				for (auto nb : g.in[i]) ufUnion(chain, rank, i, nb);
		}
		// chain tips: ordered by chain number
		std::map<size_t, std::unordered_set<size_t>> tipChainers;
		for (size_t i = 0; i < n; i++)
		{
			if (bwTip[comp[i]] || fwTip[comp[i]]) tip[i] = 1;
			if (bwTip[comp[i]])
				for (auto nb : g.out[i])
					if (chain[nb] != chain[i]) tipChainers[chain[i]].insert(nb);
			if (fwTip[comp[i]])
				for (auto nb : g.in[i])
					if (chain[nb] != chain[i]) tipChainers[chain[i]].insert(nb);
		}
		// cycles
		for (size_t i = 0; i < n; i++)
		{
			size_t fw = kNone, bw = kNone;
			for (auto u : g.out[i])
			{
				if (tip[u] || u == i) continue;
				fw = fw == kNone ? u : kNone - 1;
			}
			for (auto u : g.in[i])
			{
				if (tip[u] || u == i) continue;
				if (bw == kNone) bw = u;
				else if (u != bw) bw = kNone - 1;
			}
			if (fw != bw || fw == kNone || fw == kNone - 1) continue;
			tip[i] = 1;
			ufUnion(chain, rank, i, fw);
		}
		// bubbles, from the last piece of every bigraph node in lookup order
		for (const auto& p : g.lookup)
		{
			size_t start = p.second.back();
			auto bubble = findBubble(start, tip);
			if (!bubble.first) continue;
			std::unordered_set<size_t> visited;
			std::vector<size_t> st{start};
			ufUnion(chain, rank, start, bubble.second);
			while (!st.empty())
			{
				size_t top = st.back();
				st.pop_back();
				if (visited.count(top) || tip[top]) continue;
				visited.insert(top);
				ufUnion(chain, rank, start, top);
				for (auto nb : g.out[top])
				{
					if (visited.count(nb) || nb == bubble.second) continue;
					st.push_back(nb);
				}
			}
		}
		for (auto& p : tipChainers)
		{
			size_t unique = kNone;
			for (auto x : p.second)
			{
				if (unique == kNone) unique = chain[x];
				if (unique != chain[x])
				{
					unique = kNone - 1;
					break;
				}
			}
			if (unique == kNone - 1) continue;
			ufUnion(chain, rank, p.first, *p.second.begin());
		}
		for (size_t i = 0; i < n; i++) ufFind(chain, i);
		chainPos.assign(n, kNone);
		for (size_t s = 0; s < n; s++)
		{
			if (chainPos[s] != kNone) continue;
			size_t c = chain[s];
			std::vector<std::pair<size_t, size_t>> st{{s, (n + 5) * kSplit}};
			while (!st.empty())
			{
				auto [v, d] = st.back();
				st.pop_back();
				if (chainPos[v] != kNone) continue;
				chainPos[v] = d;
				for (auto u : g.out[v])
					if (chain[u] == c && chainPos[u] == kNone) st.emplace_back(u, d + g.len[u]);
				for (auto u : g.in[v])
					if (chain[u] == c && chainPos[u] == kNone) st.emplace_back(u, d - g.len[v]);
			}
		}
	}
};

} // namespace

char GraphHost::nodeChar(uint32_t node, uint32_t pos) const
{
	if (node < firstAmbiguous) return "ACGT"[(seq2[2 * node + pos / 32] >> ((pos % 32) * 2)) & 3];
	const uint64_t* m = &amb4[4 * (node - firstAmbiguous)];
	int set = ((m[0] >> pos) & 1) | (((m[1] >> pos) & 1) << 1) | (((m[2] >> pos) & 1) << 2) | (((m[3] >> pos) & 1) << 3);
	static const char byset[16] = {'?', 'A', 'C', 'M', 'G', 'R', 'S', 'V', 'T', 'W', 'Y', 'H', 'K', 'D', 'B', 'N'};
	return byset[set];
}

// src/AlignmentGraph.cpp:812-828
uint32_t GraphHost::unitigNode(int id, uint64_t offset) const
{
	const uint32_t* nodes;
	size_t count;
	if ((size_t)id < sizeById.size() && sizeById[(size_t)id] != 0)
	{
		nodes = byIdNodes.data() + byIdStart[(size_t)id];
		count = byIdStart[(size_t)id + 1] - byIdStart[(size_t)id];
	}
	else
	{
		const auto& v = lookup.at(id);
		nodes = v.data();
		count = v.size();
	}
	size_t index = count * ((double)offset / (double)bigSize(id));
	if (index >= count) index = count - 1;
	while (index < count - 1 && nodeOff[nodes[index]] + nodeLen[nodes[index]] <= offset) index++;
	while (index > 0 && nodeOff[nodes[index]] > offset) index--;
	return nodes[index];
}

// src/AlignmentGraph.cpp:830-847
std::pair<int, uint64_t> GraphHost::reversePosition(int id, uint64_t offset) const
{
	uint64_t size = bigSize(id);
	int rev = (id % 2 == 0) ? (id / 2) * 2 + 1 : (id / 2) * 2;
	return {rev, size - offset - 1};
}

uint64_t nextObjectUid()
{
	static std::atomic<uint64_t> next{1};
	return next.fetch_add(1);
}

GraphHost loadGfaGraph(const std::string& path)
{
	std::ifstream file{path};
	if (!file.good()) throw std::runtime_error("cannot open graph file " + path);
	Gfa gfa = parseGfa(file);
	Build b;
	// src/BigraphToDigraph.cpp:215-267
	std::unordered_map<int, std::vector<size_t>> breaks;
	for (const auto& p : gfa.overlaps)
	{
		Side fromRev{p.first.first.id, !p.first.first.plus};
		int to = p.first.second.id * 2 + (p.first.second.plus ? 0 : 1);
		int from = fromRev.id * 2 + (fromRev.plus ? 0 : 1);
		breaks[from].push_back(p.second);
		breaks[to].push_back(p.second);
	}
	for (const auto& node : gfa.nodes)
	{
		auto nameIt = gfa.names.find(node.first);
		std::string name = nameIt == gfa.names.end() ? std::string() : nameIt->second;
		std::string rc;
		rc.reserve(node.second.size());
		for (size_t i = node.second.size(); i-- > 0;) rc += complementBase(node.second[i]);
		for (char c : node.second)
			if (baseSet(c) == 0) throw std::runtime_error("invalid sequence character in graph");
		auto fw = breaks[node.first * 2];
		auto bw = breaks[node.first * 2 + 1];
		fw.push_back(0);
		fw.push_back(node.second.size());
		bw.push_back(0);
		bw.push_back(node.second.size());
		std::sort(fw.begin(), fw.end());
		std::sort(bw.begin(), bw.end());
		b.addNode(node.first * 2, node.second, name, false, fw);
		b.addNode(node.first * 2 + 1, rc, name, true, bw);
	}
	for (const auto& e : gfa.edges)
	{
		for (const auto& t : e.second)
		{
			size_t overlap = gfa.edgeOverlap;
			auto it = gfa.overlaps.find(std::make_pair(e.first, t));
			if (it != gfa.overlaps.end()) overlap = it->second;
			// src/BigraphToDigraph.cpp:106-132
			int fromRight = e.first.plus ? e.first.id * 2 : e.first.id * 2 + 1;
			int fromLeft = e.first.plus ? e.first.id * 2 + 1 : e.first.id * 2;
			int toRight = t.plus ? t.id * 2 : t.id * 2 + 1;
			int toLeft = t.plus ? t.id * 2 + 1 : t.id * 2;
			b.addEdge(fromRight, toRight, overlap);
			b.addEdge(toLeft, fromLeft, overlap);
		}
	}
	Finalizer f(b);
	size_t firstAmb = f.renumberAmbiguous();
	f.linearizable();
	f.components();
	f.chains();

	GraphHost g;
	g.uid = nextObjectUid();
	size_t n = b.len.size();
	g.firstAmbiguous = firstAmb;
	g.dbgOverlap = gfa.edgeOverlap;
	g.bpSize = b.bp;
	g.nodeLen = b.len;
	g.nodeId = b.id;
	g.nodeOff = b.off;
	g.reverse = b.rev;
	g.linearizable = f.lin;
	g.seq2.resize(2 * firstAmb);
	for (size_t i = 0; i < firstAmb; i++)
	{
		g.seq2[2 * i] = b.seq[i][0];
		g.seq2[2 * i + 1] = b.seq[i][1];
	}
	g.amb4.resize(4 * (n - firstAmb));
	for (size_t i = firstAmb; i < n; i++)
		for (int k = 0; k < 4; k++) g.amb4[4 * (i - firstAmb) + k] = b.amb[i][k];
	g.comp.assign(f.comp.begin(), f.comp.end());
	g.chain.assign(f.chain.begin(), f.chain.end());
	g.chainPos.assign(f.chainPos.begin(), f.chainPos.end());
	g.inOff.resize(n + 1);
	g.outOff.resize(n + 1);
	for (size_t i = 0; i < n; i++)
	{
		g.inOff[i] = g.inAdj.size();
		for (auto x : b.in[i]) g.inAdj.push_back(x);
		g.outOff[i] = g.outAdj.size();
		for (auto x : b.out[i]) g.outAdj.push_back(x);
	}
	g.inOff[n] = g.inAdj.size();
	g.outOff[n] = g.outAdj.size();
	for (auto& p : b.lookup)
	{
		g.lookup[p.first].assign(p.second.begin(), p.second.end());
		g.lookupOrder.push_back(p.first);
	}
	g.originalSize.insert(b.osize.begin(), b.osize.end());
	g.originalName = b.oname;
	g.splitBigSize.resize(n);
	for (size_t i = 0; i < n; i++) g.splitBigSize[i] = g.originalSize.at(g.nodeId[i]);
	{
		int64_t maxId = -1;
		bool nonNegative = true;
		for (const auto& p : g.originalSize)
		{
			maxId = std::max<int64_t>(maxId, p.first);
			nonNegative = nonNegative && p.first >= 0;
		}
		if (nonNegative && maxId >= 0 && (uint64_t)maxId < 4 * (uint64_t)g.originalSize.size() + 64)
		{
			g.sizeById.assign((size_t)maxId + 1, 0);
			g.nameById.assign((size_t)maxId + 1, 0);
			g.byIdStart.assign((size_t)maxId + 2, 0);
			for (const auto& p : g.originalSize) g.sizeById[(size_t)p.first] = p.second;
			for (const auto& p : g.originalName)
				if (p.first >= 0 && p.first <= maxId && !p.second.empty())
				{
					g.names.push_back(p.second);
					g.nameById[(size_t)p.first] = (uint32_t)g.names.size();
				}
			for (int64_t id = 0; id <= maxId; id++)
			{
				auto it = g.lookup.find((int)id);
				if (it != g.lookup.end()) g.byIdNodes.insert(g.byIdNodes.end(), it->second.begin(), it->second.end());
				g.byIdStart[(size_t)id + 1] = (uint32_t)g.byIdNodes.size();
			}
		}
	}
	g.nodeRec.resize(n);
	for (size_t i = 0; i < n; i++)
	{
		gad::NodeRec& r = g.nodeRec[i];
		const bool amb = i >= g.firstAmbiguous;
		r.seq0 = amb ? 0 : g.seq2[2 * i];
		r.seq1 = amb ? 0 : g.seq2[2 * i + 1];
		r.len = g.nodeLen[i];
		r.comp = g.comp[i];
		r.inBegin = g.inOff[i];
		r.inCount = g.inOff[i + 1] - g.inOff[i];
		r.outBegin = g.outOff[i];
		r.outCount = g.outOff[i + 1] - g.outOff[i];
		r.firstIn = r.inCount > 0 ? g.inAdj[g.inOff[i]] : ~0u;
		r.linearizable = g.linearizable[i];
	}
	g.outEdges.resize(g.outAdj.size());
	for (size_t e = 0; e < g.outAdj.size(); e++) g.outEdges[e] = gad::OutEdge{g.outAdj[e], g.comp[g.outAdj[e]]};
	return g;
}

} // namespace ga
