// Definitions behind the opaque C handles.
#pragma once
#include <memory>
#include "../../include/graphaligner_amd.h"
#include "ga_align.hpp"
#include "ga_graph.hpp"
#include "ga_index.hpp"
#include "ga_mlp.hpp"
#include "ga_proposal.hpp"
#include "ga_runtime.hpp"

struct ga_graph
{
	ga::GraphHost host;
};

struct ga_index
{
	ga::MinimizerIndex idx;
	const ga_graph* graph;
};

struct ga_context
{
	std::unique_ptr<ga::Context> ctx;
	// per-read outcomes of the last call, kept so their text buffers are reused
	std::vector<ga::ReadOutcome> outcomes;
	// true while the device read set is the one ga_reads_upload loaded; calls
	// that upload their own reads (ga_seeds, ga_align_reads, ga_align_one_way)
	// replace it and clear this, so ga_align_resident fails with GA_ESTALE
	// instead of aligning other reads under the caller's names
	bool residentFromUpload = false;
	// learned proposal applied by ga_align_reads / ga_align_resident (scorer null: off)
	ga::SeedProposal proposal;
};

struct ga_mlp
{
	std::unique_ptr<ga::MlpModel> model;
};

struct ga_records
{
	ga::ClusterRecords rec;
};

struct ga_alignment_result
{
	std::vector<int32_t> status;
	std::vector<uint64_t> seedsExtended;
	std::vector<std::vector<ga::Alignment>> alignments;
};

namespace ga {
void copyErr(char* err, size_t cap, const char* msg);
int32_t errorCostOf(const ga_params_t& p);
int32_t xdropOf(const ga_params_t& p);
gad::ExtParams extParamsOf(const ga_params_t& p);
} // namespace ga
