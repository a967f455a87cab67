// The learned proposal's network: a 3-layer perceptron over the 48 features of
// a seed cluster (48 -> 128 -> 128 -> 2, ReLU, sigmoid outputs) with BF16
// weights and activations and FP32 accumulation, trained with Adam on binary
// cross-entropy. Forward and backward run as hand-written gfx950 MFMA kernels
// (v_mfma_f32_16x16x16_bf16): records are the MFMA N dimension, so each
// layer's output tile in accumulator layout is the next layer's B operand.
//
// Parameters, flat FP32 (the layout of get/set/gradients):
//   W1[64][128] (input-major; inputs 48..63 are padding and stay 0), b1[128],
//   W2[128][128], b2[128], W3[128][16] (outputs 2..15 padding), b3[16].
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <vector>
#include "ga_proposal.hpp"

namespace gad {

constexpr int kMlpIn = 64;      // 48 features, padded
constexpr int kMlpInUsed = 48;
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 16;     // 2 outputs, padded
constexpr int kMlpOutUsed = 2;
constexpr size_t kMlpW1 = 0;
constexpr size_t kMlpB1 = kMlpW1 + (size_t)kMlpIn * kMlpHidden;
constexpr size_t kMlpW2 = kMlpB1 + kMlpHidden;
constexpr size_t kMlpB2 = kMlpW2 + (size_t)kMlpHidden * kMlpHidden;
constexpr size_t kMlpW3 = kMlpB2 + kMlpHidden;
constexpr size_t kMlpB3 = kMlpW3 + (size_t)kMlpHidden * kMlpOut;
constexpr size_t kMlpParams = kMlpB3 + kMlpOut;
// BF16 weight copies, output-major (the A operand of the forward layers):
// W1t[128][64], W2t[128][128], W3t[16][128]
constexpr size_t kMlpT1 = 0;
constexpr size_t kMlpT2 = kMlpT1 + (size_t)kMlpHidden * kMlpIn;
constexpr size_t kMlpT3 = kMlpT2 + (size_t)kMlpHidden * kMlpHidden;
constexpr size_t kMlpTWords = kMlpT3 + (size_t)kMlpOut * kMlpHidden;

// Round to nearest even (finite values).
inline __host__ __device__ uint16_t toBf16(float f)
{
	uint32_t u;
	__builtin_memcpy(&u, &f, 4);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

// Feature-major activations of one batch of bpad records (bpad a multiple of
// 16), BF16: X[64], H1[128], H2[128] after ReLU; dO[16], dH2[128], dH1[128]
// gradients of the pre-activations.
struct MlpActs
{
	uint16_t *x, *h1, *h2, *dO, *dH2, *dH1;
};

// Forward pass of records [0, nrec) of X (row-major [record][64] BF16, or
// X[idx[r]] when idx is given). probs != null: sigmoid outputs [nrec][2].
// acts.x != null: also the backward pass of the mean binary cross-entropy over
// both outputs (labels [record][2], gradient scale gradScale per output), the
// batch's activations into acts and the summed loss added to *loss.
hipError_t launchMlpForward(hipStream_t s, const uint16_t* wt, const float* params, const uint16_t* X, const uint32_t* idx, uint32_t nrec, uint32_t bpad, float* probs, const float* labels, float gradScale, MlpActs acts, float* loss);
// Weight and bias gradients of a batch from its activations: grads (flat FP32).
hipError_t launchMlpWeightGrads(hipStream_t s, MlpActs acts, uint32_t bpad, float* grads);
// One Adam step (step t >= 1) on params with grads, then the BF16 copies; with
// lr = 0 only the copies are rewritten.
hipError_t launchMlpAdam(hipStream_t s, float* params, const float* grads, float* m, float* v, uint32_t t, float lr, float beta1, float beta2, float eps, uint16_t* wt);

} // namespace gad

namespace ga {

struct MlpTrainOptions
{
	uint32_t epochs = 4;
	uint32_t batch = 4096;
	float lr = 2e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
	uint64_t seed = 1;
};

struct MlpTrainStats
{
	std::vector<float> epochLoss;  // mean loss over each epoch's batches
	double ms = 0;                 // device time of the training launches
	uint64_t steps = 0;
};

// The network on one GPU (device memory and a stream of its own). Calls on one
// model are serialised by its mutex.
class MlpModel : public ClusterScorer
{
public:
	MlpModel(int device, uint64_t seed);
	~MlpModel() override;
	MlpModel(const MlpModel&) = delete;
	MlpModel& operator=(const MlpModel&) = delete;
	void getWeights(float* out);
	void setWeights(const float* w);
	// n rows of kMlpInUsed features (FP32, rounded to BF16) -> n rows of 2 probabilities
	void forward(const float* features, size_t n, float* out);
	// mean binary cross-entropy of n records and its gradient (flat layout)
	double gradients(const float* features, const float* labels, size_t n, float* grads);
	MlpTrainStats train(const float* features, const float* labels, size_t n, const MlpTrainOptions& opt);
	void score(const float* features, size_t n, float* out) override { forward(features, n, out); }
	int device() const { return device_; }

private:
	void ensureActs(uint32_t bpad);
	void uploadFeatures(const float* features, size_t n);
	int device_;
	std::mutex mu_;
	hipStream_t stream_ = nullptr;
	float *params_ = nullptr, *grads_ = nullptr, *m_ = nullptr, *v_ = nullptr, *loss_ = nullptr;
	uint16_t* wt_ = nullptr;
	uint16_t* x_ = nullptr;
	size_t xCap_ = 0;
	float* labels_ = nullptr;
	size_t labelCap_ = 0;
	float* probs_ = nullptr;
	size_t probCap_ = 0;
	uint32_t* idx_ = nullptr;
	size_t idxCap_ = 0;
	uint16_t* acts_ = nullptr;
	uint32_t actsCap_ = 0;
	uint32_t adamStep_ = 0;
	std::vector<uint16_t> hostX_;
};

} // namespace ga
