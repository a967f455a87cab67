// gfx950 kernels of the proposal network and its host-side model.
//
// MFMA v_mfma_f32_16x16x16_bf16 (one wave): lane l = 16 q + r (r = l % 16,
// q = l / 16) supplies A[r][4q .. 4q+3] and B[4q .. 4q+3][r] and holds
// D[4q .. 4q+3][r]. A layer computes H_next^T = W^T H^T with the layer's output
// features as M and a tile of 16 records as N, so the accumulator tile of
// output features 16m .. 16m+15 is, after bias, ReLU and rounding to BF16,
// exactly the B operand of the next layer's K-step m: activations never leave
// registers between layers. The backward pass runs the same way with W in
// place of W^T (read transposed from the output-major copy in LDS).
//
// Weight gradients dW = sum over records of H dH^T are products whose K
// dimension is the batch: the forward kernel stores every activation
// feature-major ([feature][record]), where the A operand (4 consecutive
// records of one input feature) and the B operand (4 consecutive records of
// one output feature) are single 8-byte loads; one wave owns one 16x16 tile of
// a weight matrix and runs its whole K loop, so no atomics and no partial sums
// are needed and the result does not depend on the launch shape.
#include "ga_mlp.hpp"
#include "ga_runtime.hpp"

#include <chrono>
#include <cmath>

namespace gad {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline bf16x4 packBf16(f32x4 v)
{
	bf16x4 r;
	r.x = (short)toBf16(v.x);
	r.y = (short)toBf16(v.y);
	r.z = (short)toBf16(v.z);
	r.w = (short)toBf16(v.w);
	return r;
}

__device__ inline float sigmoidOf(float z) { return 1.0f / (1.0f + expf(-z)); }

// LDS rows of the output-major BF16 weights, padded by 4 elements (8 bytes) so
// that the 16 rows one A operand spans start in different banks
constexpr int kS1 = kMlpIn + 4;
constexpr int kS2 = kMlpHidden + 4;
constexpr int kS3 = kMlpHidden + 4;

struct MlpLds
{
	uint16_t w1t[kMlpHidden * kS1];
	uint16_t w2t[kMlpHidden * kS2];
	uint16_t w3t[kMlpOut * kS3];
	float b1[kMlpHidden], b2[kMlpHidden], b3[kMlpOut];
};
static_assert(sizeof(MlpLds) <= 64 * 1024, "proposal network weights exceed one workgroup's LDS");

__device__ inline void stageRows(uint16_t* dst, int stride, const uint16_t* src, int rows, int cols)
{
	const int chunks = cols / 4;
	for (int e = threadIdx.x; e < rows * chunks; e += blockDim.x)
	{
		const int r = e / chunks, c = e % chunks;
		*(uint64_t*)(dst + r * stride + c * 4) = *(const uint64_t*)(src + (size_t)r * cols + c * 4);
	}
}

template <bool Train>
__global__ void __launch_bounds__(256) mlpForwardKernel(const uint16_t* __restrict__ wt, const float* __restrict__ params, const uint16_t* __restrict__ X, const uint32_t* __restrict__ idx, uint32_t nrec, uint32_t bpad, float* __restrict__ probs, const float* __restrict__ labels, float gradScale, MlpActs acts, float* __restrict__ loss)
{
	__shared__ MlpLds L;
	stageRows(L.w1t, kS1, wt + kMlpT1, kMlpHidden, kMlpIn);
	stageRows(L.w2t, kS2, wt + kMlpT2, kMlpHidden, kMlpHidden);
	stageRows(L.w3t, kS3, wt + kMlpT3, kMlpOut, kMlpHidden);
	for (int e = threadIdx.x; e < kMlpHidden; e += blockDim.x)
	{
		L.b1[e] = params[kMlpB1 + e];
		L.b2[e] = params[kMlpB2 + e];
		if (e < kMlpOut) L.b3[e] = params[kMlpB3 + e];
	}
	__syncthreads();
	const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t r = lane & 15, q = lane >> 4;
	const uint32_t tiles = bpad / 16;
	const f32x4 zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
	const bf16x4 zeroB = {0, 0, 0, 0};
	for (uint32_t t = blockIdx.x * 4 + wave; t < tiles; t += gridDim.x * 4)
	{
		const uint32_t col = t * 16 + r;
		const bool valid = col < nrec;
		const uint32_t rec = valid ? (idx ? idx[col] : col) : 0;
		bf16x4 xb[kMlpIn / 16];
		for (int ks = 0; ks < kMlpIn / 16; ks++) xb[ks] = valid ? *(const bf16x4*)(X + (size_t)rec * kMlpIn + ks * 16 + q * 4) : zeroB;
		f32x4 h1[8], h2[8];
		bf16x4 h1b[8], h2b[8];
		for (int m = 0; m < 8; m++)
		{
			f32x4 acc = zero4;
			for (int ks = 0; ks < kMlpIn / 16; ks++) acc = mfma16(*(const bf16x4*)(L.w1t + (m * 16 + r) * kS1 + ks * 16 + q * 4), xb[ks], acc);
			for (int a = 0; a < 4; a++) acc[a] = fmaxf(acc[a] + L.b1[m * 16 + q * 4 + a], 0.0f);
			h1[m] = acc;
			h1b[m] = packBf16(acc);
		}
		for (int m = 0; m < 8; m++)
		{
			f32x4 acc = zero4;
			for (int ks = 0; ks < 8; ks++) acc = mfma16(*(const bf16x4*)(L.w2t + (m * 16 + r) * kS2 + ks * 16 + q * 4), h1b[ks], acc);
			for (int a = 0; a < 4; a++) acc[a] = fmaxf(acc[a] + L.b2[m * 16 + q * 4 + a], 0.0f);
			h2[m] = acc;
			h2b[m] = packBf16(acc);
		}
		f32x4 o = zero4;
		for (int ks = 0; ks < 8; ks++) o = mfma16(*(const bf16x4*)(L.w3t + r * kS3 + ks * 16 + q * 4), h2b[ks], o);
		for (int a = 0; a < 4; a++) o[a] += L.b3[q * 4 + a];
		if (!Train)
		{
			if (q == 0 && valid)
			{
				probs[(size_t)col * 2 + 0] = sigmoidOf(o[0]);
				probs[(size_t)col * 2 + 1] = sigmoidOf(o[1]);
			}
			continue;
		}
		// binary cross-entropy with logits on outputs 0 and 1 (lanes q = 0)
		float lossSum = 0.0f;
		f32x4 g = zero4;
		if (q == 0 && valid)
			for (int a = 0; a < kMlpOutUsed; a++)
			{
				const float z = o[a], y = labels[(size_t)rec * 2 + a];
				lossSum += fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
				g[a] = (sigmoidOf(z) - y) * gradScale;
			}
		const bf16x4 gb = packBf16(g);
		// dH2^T = W3 dO^T, masked by ReLU'
		bf16x4 dh2b[8];
		for (int m = 0; m < 8; m++)
		{
			bf16x4 w;
			for (int j = 0; j < 4; j++) w[j] = (short)L.w3t[(q * 4 + j) * kS3 + m * 16 + r];
			f32x4 acc = mfma16(w, gb, zero4);
			for (int a = 0; a < 4; a++) acc[a] = h2[m][a] > 0.0f ? acc[a] : 0.0f;
			dh2b[m] = packBf16(acc);
		}
		// dH1^T = W2 dH2^T, masked by ReLU'
		bf16x4 dh1b[8];
		for (int m = 0; m < 8; m++)
		{
			f32x4 acc = zero4;
			for (int ks = 0; ks < 8; ks++)
			{
				bf16x4 w;
				for (int j = 0; j < 4; j++) w[j] = (short)L.w2t[(ks * 16 + q * 4 + j) * kS2 + m * 16 + r];
				acc = mfma16(w, dh2b[ks], acc);
			}
			for (int a = 0; a < 4; a++) acc[a] = h1[m][a] > 0.0f ? acc[a] : 0.0f;
			dh1b[m] = packBf16(acc);
		}
		// feature-major activations of the batch (every column < bpad is written)
		for (int a = 0; a < 4; a++)
		{
			for (int ks = 0; ks < kMlpIn / 16; ks++) acts.x[(size_t)(ks * 16 + q * 4 + a) * bpad + col] = (uint16_t)xb[ks][a];
			for (int m = 0; m < 8; m++)
			{
				const size_t row = (size_t)(m * 16 + q * 4 + a) * bpad + col;
				acts.h1[row] = (uint16_t)h1b[m][a];
				acts.h2[row] = (uint16_t)h2b[m][a];
				acts.dH2[row] = (uint16_t)dh2b[m][a];
				acts.dH1[row] = (uint16_t)dh1b[m][a];
			}
			acts.dO[(size_t)(q * 4 + a) * bpad + col] = (uint16_t)gb[a];
		}
		for (int off = 32; off > 0; off >>= 1) lossSum += __shfl_xor(lossSum, off);
		if (lane == 0 && lossSum != 0.0f) atomicAdd(loss, lossSum);
	}
}

// One wave per 16x16 tile of dW1 (4x8 tiles), dW2 (8x8), dW3 (8x1), then one
// wave per 16 bias rows (b1, b2: 8 each; b3: 1).
constexpr uint32_t kGradTilesW1 = 4 * 8, kGradTilesW2 = 8 * 8, kGradTilesW3 = 8;
constexpr uint32_t kGradTiles = kGradTilesW1 + kGradTilesW2 + kGradTilesW3;
constexpr uint32_t kBiasWaves = 8 + 8 + 1;

__global__ void __launch_bounds__(64) mlpWeightGradKernel(MlpActs acts, uint32_t bpad, float* __restrict__ grads)
{
	const uint32_t lane = threadIdx.x, r = lane & 15, q = lane >> 4;
	uint32_t w = blockIdx.x;
	if (w < kGradTiles)
	{
		const uint16_t *A, *D;
		uint32_t ib, jb, outDim;
		size_t off;
		if (w < kGradTilesW1)
		{
			A = acts.x, D = acts.dH1, ib = w / 8, jb = w % 8, outDim = kMlpHidden, off = kMlpW1;
		}
		else if (w < kGradTilesW1 + kGradTilesW2)
		{
			w -= kGradTilesW1;
			A = acts.h1, D = acts.dH2, ib = w / 8, jb = w % 8, outDim = kMlpHidden, off = kMlpW2;
		}
		else
		{
			w -= kGradTilesW1 + kGradTilesW2;
			A = acts.h2, D = acts.dO, ib = w, jb = 0, outDim = kMlpOut, off = kMlpW3;
		}
		const uint16_t* a = A + (size_t)(ib * 16 + r) * bpad + q * 4;
		const uint16_t* d = D + (size_t)(jb * 16 + r) * bpad + q * 4;
		f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
		for (uint32_t kb = 0; kb < bpad; kb += 16) acc = mfma16(*(const bf16x4*)(a + kb), *(const bf16x4*)(d + kb), acc);
		for (int j = 0; j < 4; j++) grads[off + (size_t)(ib * 16 + q * 4 + j) * outDim + jb * 16 + r] = acc[j];
		return;
	}
	w -= kGradTiles;
	if (w >= kBiasWaves) return;
	const uint16_t* D;
	uint32_t row;
	size_t off;
	if (w < 8) D = acts.dH1, row = w * 16 + r, off = kMlpB1;
	else if (w < 16) D = acts.dH2, row = (w - 8) * 16 + r, off = kMlpB2;
	else D = acts.dO, row = r, off = kMlpB3;
	float sum = 0.0f;
	const uint16_t* d = D + (size_t)row * bpad + q * 4;
	for (uint32_t kb = 0; kb < bpad; kb += 16)
	{
		const bf16x4 v = *(const bf16x4*)(d + kb);
		for (int j = 0; j < 4; j++) sum += __uint_as_float((uint32_t)(uint16_t)v[j] << 16);
	}
	sum += __shfl_xor(sum, 16);
	sum += __shfl_xor(sum, 32);
	if (q == 0) grads[off + row] = sum;
}

__global__ void __launch_bounds__(256) mlpAdamKernel(float* __restrict__ params, const float* __restrict__ grads, float* __restrict__ m, float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1, float c2, uint16_t* __restrict__ wt)
{
	const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= kMlpParams) return;
	float w = params[p];
	if (lr > 0.0f)
	{
		const float g = grads[p];
		const float mp = beta1 * m[p] + (1.0f - beta1) * g;
		const float vp = beta2 * v[p] + (1.0f - beta2) * g * g;
		m[p] = mp;
		v[p] = vp;
		w -= lr * (mp * c1) / (sqrtf(vp * c2) + eps);
		params[p] = w;
	}
	if (p < kMlpB1)
		wt[kMlpT1 + (p % kMlpHidden) * kMlpIn + p / kMlpHidden] = toBf16(w);
	else if (p >= kMlpW2 && p < kMlpB2)
	{
		const size_t e = p - kMlpW2;
		wt[kMlpT2 + (e % kMlpHidden) * kMlpHidden + e / kMlpHidden] = toBf16(w);
	}
	else if (p >= kMlpW3 && p < kMlpB3)
	{
		const size_t e = p - kMlpW3;
		wt[kMlpT3 + (e % kMlpOut) * kMlpHidden + e / kMlpOut] = toBf16(w);
	}
}

__global__ void __launch_bounds__(64) fillFloatKernel(float* p, uint32_t n)
{
	for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0f;
}

} // namespace

hipError_t launchMlpForward(hipStream_t s, const uint16_t* wt, const float* params, const uint16_t* X, const uint32_t* idx, uint32_t nrec, uint32_t bpad, float* probs, const float* labels, float gradScale, MlpActs acts, float* loss)
{
	if (bpad % 16 || nrec > bpad) return hipErrorInvalidValue;
	if (bpad == 0) return hipSuccess;
	const uint32_t tiles = bpad / 16;
	// four 16-record tiles per workgroup, at most 1024 workgroups looping
	const uint32_t blocks = std::min<uint32_t>((tiles + 3) / 4, 1024);
	if (acts.x)
		hipLaunchKernelGGL(mlpForwardKernel<true>, dim3(blocks), dim3(256), 0, s, wt, params, X, idx, nrec, bpad, probs, labels, gradScale, acts, loss);
	else
		hipLaunchKernelGGL(mlpForwardKernel<false>, dim3(blocks), dim3(256), 0, s, wt, params, X, idx, nrec, bpad, probs, labels, gradScale, acts, loss);
	return hipGetLastError();
}

hipError_t launchMlpWeightGrads(hipStream_t s, MlpActs acts, uint32_t bpad, float* grads)
{
	if (bpad % 16 || bpad == 0) return hipErrorInvalidValue;
	hipLaunchKernelGGL(mlpWeightGradKernel, dim3(kGradTiles + kBiasWaves), dim3(64), 0, s, acts, bpad, grads);
	return hipGetLastError();
}

hipError_t launchMlpAdam(hipStream_t s, float* params, const float* grads, float* m, float* v, uint32_t t, float lr, float beta1, float beta2, float eps, uint16_t* wt)
{
	const float c1 = t ? (float)(1.0 / (1.0 - std::pow((double)beta1, (double)t))) : 1.0f;
	const float c2 = t ? (float)(1.0 / (1.0 - std::pow((double)beta2, (double)t))) : 1.0f;
	hipLaunchKernelGGL(mlpAdamKernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, grads, m, v, lr, beta1, beta2, eps, c1, c2, wt);
	return hipGetLastError();
}

static hipError_t zeroFloats(hipStream_t s, float* p, uint32_t n)
{
	hipLaunchKernelGGL(fillFloatKernel, dim3(1), dim3(64), 0, s, p, n);
	return hipGetLastError();
}

} // namespace gad

namespace ga {

namespace {

uint64_t splitmix(uint64_t& x)
{
	uint64_t z = (x += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

double uniform(uint64_t& state) { return (double)(splitmix(state) >> 11) * (1.0 / 9007199254740992.0); }

uint32_t roundUp16(size_t n) { return (uint32_t)((n + 15) / 16 * 16); }

template <class T>
void growDevice(T*& p, size_t& cap, size_t n)
{
	if (n <= cap) return;
	if (p) GA_HIP(hipFree(p));
	p = nullptr;
	cap = 0;
	GA_HIP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
	cap = n;
}

} // namespace

MlpModel::MlpModel(int device, uint64_t seed) : device_(device)
{
	GA_HIP(hipSetDevice(device_));
	GA_HIP(hipStreamCreate(&stream_));
	GA_HIP(hipMalloc((void**)&params_, gad::kMlpParams * sizeof(float)));
	GA_HIP(hipMalloc((void**)&grads_, gad::kMlpParams * sizeof(float)));
	GA_HIP(hipMalloc((void**)&m_, gad::kMlpParams * sizeof(float)));
	GA_HIP(hipMalloc((void**)&v_, gad::kMlpParams * sizeof(float)));
	GA_HIP(hipMalloc((void**)&loss_, 64 * sizeof(float)));
	GA_HIP(hipMalloc((void**)&wt_, gad::kMlpTWords * sizeof(uint16_t)));
	// He-uniform hidden layers, Glorot-uniform output layer, zero biases and padding
	std::vector<float> w(gad::kMlpParams, 0.0f);
	uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
	const double a1 = std::sqrt(6.0 / gad::kMlpInUsed), a2 = std::sqrt(6.0 / gad::kMlpHidden), a3 = std::sqrt(6.0 / (gad::kMlpHidden + gad::kMlpOutUsed));
	for (int i = 0; i < gad::kMlpInUsed; i++)
		for (int j = 0; j < gad::kMlpHidden; j++) w[gad::kMlpW1 + (size_t)i * gad::kMlpHidden + j] = (float)((2 * uniform(st) - 1) * a1);
	for (int i = 0; i < gad::kMlpHidden; i++)
		for (int j = 0; j < gad::kMlpHidden; j++) w[gad::kMlpW2 + (size_t)i * gad::kMlpHidden + j] = (float)((2 * uniform(st) - 1) * a2);
	for (int i = 0; i < gad::kMlpHidden; i++)
		for (int j = 0; j < gad::kMlpOutUsed; j++) w[gad::kMlpW3 + (size_t)i * gad::kMlpOut + j] = (float)((2 * uniform(st) - 1) * a3);
	setWeights(w.data());
}

MlpModel::~MlpModel()
{
	(void)hipSetDevice(device_);
	if (stream_) (void)hipStreamSynchronize(stream_);
	for (void* p : {(void*)params_, (void*)grads_, (void*)m_, (void*)v_, (void*)loss_, (void*)wt_, (void*)x_, (void*)labels_, (void*)probs_, (void*)idx_, (void*)acts_})
		if (p) (void)hipFree(p);
	if (stream_) (void)hipStreamDestroy(stream_);
}

void MlpModel::getWeights(float* out)
{
	std::lock_guard<std::mutex> lock(mu_);
	GA_HIP(hipSetDevice(device_));
	GA_HIP(hipMemcpyAsync(out, params_, gad::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
}

void MlpModel::setWeights(const float* w)
{
	std::lock_guard<std::mutex> lock(mu_);
	GA_HIP(hipSetDevice(device_));
	GA_HIP(hipMemcpyAsync(params_, w, gad::kMlpParams * sizeof(float), hipMemcpyHostToDevice, stream_));
	GA_HIP(hipMemsetAsync(m_, 0, gad::kMlpParams * sizeof(float), stream_));
	GA_HIP(hipMemsetAsync(v_, 0, gad::kMlpParams * sizeof(float), stream_));
	adamStep_ = 0;
	GA_HIP(gad::launchMlpAdam(stream_, params_, grads_, m_, v_, 0, 0.0f, 0.9f, 0.999f, 1e-8f, wt_));
	GA_HIP(hipStreamSynchronize(stream_));
}

void MlpModel::uploadFeatures(const float* features, size_t n)
{
	hostX_.assign(n * gad::kMlpIn, 0);
	for (size_t i = 0; i < n; i++)
		for (int f = 0; f < gad::kMlpInUsed; f++) hostX_[i * gad::kMlpIn + f] = gad::toBf16(features[i * gad::kMlpInUsed + f]);
	growDevice(x_, xCap_, hostX_.size());
	if (n) GA_HIP(hipMemcpyAsync(x_, hostX_.data(), hostX_.size() * sizeof(uint16_t), hipMemcpyHostToDevice, stream_));
}

void MlpModel::ensureActs(uint32_t bpad)
{
	if (bpad <= actsCap_) return;
	if (acts_) GA_HIP(hipFree(acts_));
	acts_ = nullptr;
	actsCap_ = 0;
	const size_t rows = gad::kMlpIn + 4 * gad::kMlpHidden + gad::kMlpOut;
	GA_HIP(hipMalloc((void**)&acts_, rows * bpad * sizeof(uint16_t)));
	actsCap_ = bpad;
}

static gad::MlpActs actsView(uint16_t* base, uint32_t bpad)
{
	gad::MlpActs a;
	a.x = base;
	a.h1 = a.x + (size_t)gad::kMlpIn * bpad;
	a.h2 = a.h1 + (size_t)gad::kMlpHidden * bpad;
	a.dH2 = a.h2 + (size_t)gad::kMlpHidden * bpad;
	a.dH1 = a.dH2 + (size_t)gad::kMlpHidden * bpad;
	a.dO = a.dH1 + (size_t)gad::kMlpHidden * bpad;
	return a;
}

void MlpModel::forward(const float* features, size_t n, float* out)
{
	std::lock_guard<std::mutex> lock(mu_);
	if (n == 0) return;
	if (n > (1u << 30)) throw std::runtime_error("too many records for one forward pass");
	GA_HIP(hipSetDevice(device_));
	uploadFeatures(features, n);
	growDevice(probs_, probCap_, n * 2);
	GA_HIP(gad::launchMlpForward(stream_, wt_, params_, x_, nullptr, (uint32_t)n, roundUp16(n), probs_, nullptr, 0.0f, gad::MlpActs{}, nullptr));
	GA_HIP(hipMemcpyAsync(out, probs_, n * 2 * sizeof(float), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
}

double MlpModel::gradients(const float* features, const float* labels, size_t n, float* grads)
{
	std::lock_guard<std::mutex> lock(mu_);
	if (n == 0) throw std::runtime_error("no records");
	if (n > (1u << 24)) throw std::runtime_error("too many records for one gradient");
	GA_HIP(hipSetDevice(device_));
	uploadFeatures(features, n);
	growDevice(labels_, labelCap_, n * 2);
	GA_HIP(hipMemcpyAsync(labels_, labels, n * 2 * sizeof(float), hipMemcpyHostToDevice, stream_));
	const uint32_t bpad = roundUp16(n);
	ensureActs(bpad);
	GA_HIP(gad::zeroFloats(stream_, loss_, 1));
	GA_HIP(gad::launchMlpForward(stream_, wt_, params_, x_, nullptr, (uint32_t)n, bpad, nullptr, labels_, (float)(1.0 / (2.0 * n)), actsView(acts_, bpad), loss_));
	GA_HIP(gad::launchMlpWeightGrads(stream_, actsView(acts_, bpad), bpad, grads_));
	float loss = 0;
	GA_HIP(hipMemcpyAsync(grads, grads_, gad::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipMemcpyAsync(&loss, loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
	GA_HIP(hipStreamSynchronize(stream_));
	return loss / (2.0 * n);
}

MlpTrainStats MlpModel::train(const float* features, const float* labels, size_t n, const MlpTrainOptions& opt)
{
	std::lock_guard<std::mutex> lock(mu_);
	MlpTrainStats out;
	if (n == 0 || opt.epochs == 0) return out;
	if (opt.batch == 0 || opt.batch > (1u << 20)) throw std::runtime_error("batch size must be 1 to 2^20");
	if (opt.epochs > 64) throw std::runtime_error("at most 64 epochs per call");
	if (n > (1u << 30)) throw std::runtime_error("too many records");
	GA_HIP(hipSetDevice(device_));
	uploadFeatures(features, n);
	growDevice(labels_, labelCap_, n * 2);
	GA_HIP(hipMemcpyAsync(labels_, labels, n * 2 * sizeof(float), hipMemcpyHostToDevice, stream_));
	growDevice(idx_, idxCap_, n);
	const uint32_t maxPad = roundUp16(std::min<size_t>(n, opt.batch));
	ensureActs(maxPad);
	GA_HIP(gad::zeroFloats(stream_, loss_, opt.epochs));
	hipEvent_t e0, e1;
	GA_HIP(hipEventCreate(&e0));
	GA_HIP(hipEventCreate(&e1));
	std::vector<uint32_t> perm(n);
	uint64_t st = opt.seed * 0x9E3779B97F4A7C15ull + 7;
	float ms = 0;
	try
	{
		for (uint32_t ep = 0; ep < opt.epochs; ep++)
		{
			for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
			for (size_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(st) % (i + 1)]);
			// the previous epoch's kernels read idx_: wait before overwriting it
			GA_HIP(hipStreamSynchronize(stream_));
			GA_HIP(hipMemcpyAsync(idx_, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
			GA_HIP(hipEventRecord(e0, stream_));
			for (size_t b0 = 0; b0 < n; b0 += opt.batch)
			{
				const uint32_t nrec = (uint32_t)std::min<size_t>(opt.batch, n - b0);
				const uint32_t bpad = roundUp16(nrec);
				const gad::MlpActs acts = actsView(acts_, bpad);
				GA_HIP(gad::launchMlpForward(stream_, wt_, params_, x_, idx_ + b0, nrec, bpad, nullptr, labels_, (float)(1.0 / (2.0 * nrec)), acts, loss_ + ep));
				GA_HIP(gad::launchMlpWeightGrads(stream_, acts, bpad, grads_));
				adamStep_++;
				GA_HIP(gad::launchMlpAdam(stream_, params_, grads_, m_, v_, adamStep_, opt.lr, opt.beta1, opt.beta2, opt.eps, wt_));
				out.steps++;
			}
			GA_HIP(hipEventRecord(e1, stream_));
			GA_HIP(hipEventSynchronize(e1));
			float epMs = 0;
			GA_HIP(hipEventElapsedTime(&epMs, e0, e1));
			ms += epMs;
		}
		std::vector<float> loss(opt.epochs);
		GA_HIP(hipMemcpyAsync(loss.data(), loss_, opt.epochs * sizeof(float), hipMemcpyDeviceToHost, stream_));
		GA_HIP(hipStreamSynchronize(stream_));
		// each batch's loss is a sum over its records: the epoch mean is the sum
		// over all records over 2 n
		for (float l : loss) out.epochLoss.push_back((float)(l / (2.0 * n)));
	}
	catch (...)
	{
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		throw;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	out.ms = ms;
	return out;
}

} // namespace ga
