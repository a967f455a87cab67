// Host-side alignment graph: GFA -> bidirected-to-directed graph -> split nodes,
// component order, linearizable flags and chains, laid out as flat arrays (CSR
// adjacency, 2-bit packed node sequences) ready to be copied into HBM.
//
// Follows src/GfaGraph.cpp:170-396, src/BigraphToDigraph.cpp:101-132,215-267 and
// src/AlignmentGraph.cpp (node numbering, IUPAC renumbering, Tarjan order,
// linearizable detection, chains), keeping the same standard-container iteration
// orders because node numbering depends on them.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>
#include "ga_device_types.h"

namespace ga {

// Process-unique, never reused identity of a loaded graph or built index (device
// copies are cached by it, so a new object at a freed one's address is not
// mistaken for it).
uint64_t nextObjectUid();

struct GraphHost
{
	uint64_t uid = 0;
	// split nodes (<= 64 bp), indexed 0..N-1; ambiguous (IUPAC) nodes are numbered last
	std::vector<uint32_t> nodeLen;
	std::vector<int32_t> nodeId;       // bigraph node id: 2*id (+) or 2*id+1 (-)
	std::vector<uint32_t> nodeOff;     // offset of the split node inside its bigraph node
	std::vector<uint8_t> reverse;
	std::vector<uint8_t> linearizable;
	std::vector<uint64_t> seq2;        // two 32-bp chunks per non-ambiguous node, A0 C1 G2 T3, LSB first
	std::vector<uint64_t> amb4;        // A,C,G,T masks per ambiguous node (index node - firstAmbiguous)
	std::vector<uint32_t> comp;        // strongly connected component number (topological)
	std::vector<uint32_t> chain;
	std::vector<uint64_t> chainPos;
	std::vector<uint32_t> inOff, inAdj, outOff, outAdj;
	uint32_t firstAmbiguous = 0;
	uint64_t dbgOverlap = 0;
	uint64_t bpSize = 0;
	// bigraph nodes
	std::unordered_map<int, std::vector<uint32_t>> lookup; // bigraph id -> split nodes in offset order
	std::unordered_map<int, uint64_t> originalSize;
	std::unordered_map<int, std::string> originalName;
	std::vector<int> lookupOrder;                         // bigraph ids in the lookup map's iteration order
	std::vector<uint64_t> splitBigSize;                   // per split node: length of its bigraph node
	// the maps above as arrays indexed by bigraph id, filled when the ids are
	// dense (below 4 x the number of bigraph nodes + 64), else empty: lengths (0
	// for an id without a node), names (index + 1 into names, 0 for none or an
	// unnamed node) and split-node lists (offsets into byIdNodes)
	std::vector<uint64_t> sizeById;
	std::vector<uint32_t> nameById;
	std::vector<std::string> names;
	std::vector<uint32_t> byIdStart, byIdNodes;
	// originalSize.at(id), originalName.at(id) through the arrays when filled
	uint64_t bigSize(int id) const
	{
		if ((size_t)id < sizeById.size() && sizeById[(size_t)id] != 0) return sizeById[(size_t)id];
		return originalSize.at(id);
	}
	// the node's name, or null for a node named by its number
	const std::string* bigName(int id) const
	{
		if ((size_t)id < nameById.size())
		{
			const uint32_t k = nameById[(size_t)id];
			return k == 0 ? nullptr : &names[k - 1];
		}
		const std::string& nm = originalName.at(id);
		return nm.empty() ? nullptr : &nm;
	}
	// the same per-node data as one record per node and the out-edges with their
	// targets' components (gad::NodeRec, gad::OutEdge), for the device layout
	std::vector<gad::NodeRec> nodeRec;
	std::vector<gad::OutEdge> outEdges;

	size_t numNodes() const { return nodeLen.size(); }
	char nodeChar(uint32_t node, uint32_t pos) const;
	uint32_t unitigNode(int id, uint64_t offset) const;
	std::pair<int, uint64_t> reversePosition(int id, uint64_t offset) const;
};

// Throws std::runtime_error on malformed input.
GraphHost loadGfaGraph(const std::string& path);

} // namespace ga
