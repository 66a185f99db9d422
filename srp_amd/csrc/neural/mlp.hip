/* Training, inference and evaluation of the neural fragment shader's
 * network (srp/srp_neural.h, layouts in srp/neural_abi.h).
 *
 * One optimiser step is three launches:
 *  1. k_train_pass: one wave per 64 samples.  Forward on the matrix cores
 *     (transposed layers, srp/device/neural.hpp), the loss, and the
 *     backward pass down to the first layer's pre-activation gradient.  The
 *     operands the weight gradients need are stored transposed, sample
 *     index contiguous, as bf16: encoded inputs X, hidden activations H1,
 *     H2 and pre-activation gradients dZ1, dZ2, dY.
 *  2. k_weight_grads: every weight and bias gradient as a product over the
 *     batch, dW = dZ^T-rows x activations (bias: x ones), split into
 *     1024-sample chunks; each chunk writes its partial sums (no atomics,
 *     deterministic).
 *  3. k_adam: sums the chunks, updates the fp32 master weights with Adam
 *     and rewrites the bf16 matrix-core copies. */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "kernels.h"
#include "srp/device/neural.hpp"
#include "srp/srp_neural.h"

using namespace srp::nn;
using srp::lib::libMessage;
using srp::lib::libStream;

#define NN_CHECK(x)                                                                         \
	do {                                                                                    \
		hipError_t e_ = (x);                                                                \
		if (e_ != hipSuccess)                                                               \
		{                                                                                   \
			fprintf(stderr, "libsrp_amd: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
			        __FILE__, __LINE__);                                                    \
			abort();                                                                        \
		}                                                                                   \
	} while (0)

namespace {

constexpr int kChunk = 1024;        /* samples per weight-gradient partial */
constexpr int kMaxBatch = 65536;
constexpr int kMaxChunks = kMaxBatch / kChunk;

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ inline float bfRound(float v) { return (float) (__bf16) v; }

/* ------------------------------------------------------------------ kernels */

struct TrainArgs
{
	const float* feat;       /* [n][8] */
	const float* targ;       /* [n][3] */
	uint64_t n;
	uint32_t sequential;     /* 1: sample s = dataset row s */
	uint32_t step;
	uint64_t seed;
	const SrpNnDevice* net;
	uint32_t batch;
	__bf16* xt;              /* [512][batch] */
	__bf16* h1t;             /* [64][batch] */
	__bf16* h2t;
	__bf16* dz1t;
	__bf16* dz2t;
	__bf16* dyt;             /* [16][batch], rows 3..15 stay zero */
	float* lossPart;         /* [batch/64] */
};

/* Store a 64 x 64 tile (rows 16mb+4g+r, columns 16nb+c of this wave's
 * samples) transposed: dst[row * ld + col]. */
__device__ inline void storeTile(__bf16* dst, size_t ld, const f32x4 t[4][4], int g, int c, bool relu)
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
#pragma unroll
			for (int r = 0; r < 4; r++)
			{
				float v = t[mb][nb][r];
				if (relu) v = v > 0.0f ? v : 0.0f;
				dst[(size_t) (16 * mb + 4 * g + r) * ld + 16 * nb + c] = (__bf16) v;
			}
}

__global__ __launch_bounds__(64) void k_train_pass(TrainArgs A)
{
	const int lane = (int) threadIdx.x;
	const int g = lane >> 4, c = lane & 15;
	const size_t s0 = (size_t) blockIdx.x * 64;
	const size_t s = s0 + (size_t) lane;
	const uint64_t row = A.sequential ? (uint64_t) s
	                                  : mix64(A.seed + 0x9E3779B97F4A7C15ull * ((uint64_t) A.step * A.batch + s + 1)) % A.n;
	float xin[SRP_NN_IN], t[3];
#pragma unroll
	for (int d = 0; d < SRP_NN_IN; d++)
		xin[d] = A.feat[row * SRP_NN_IN + d];
#pragma unroll
	for (int ch = 0; ch < 3; ch++)
		t[ch] = A.targ[row * 3 + ch];
	const SrpNnDevice* net = A.net;
	const size_t ld = A.batch;

	/* forward */
	f32x4 z1[4][4];
	layer1(net->w1, net->step, xin, g, c, z1, A.xt + s0, ld);
	biasInPlace(z1, net->b1, g);
	storeTile(A.h1t + s0, ld, z1, g, c, true);
	bf16x8 h[2][4];
	packOperands<true>(z1, h);
	f32x4 z2[4][4];
	layerHidden(net->w2f, h, g, c, z2);
	biasInPlace(z2, net->b2, g);
	storeTile(A.h2t + s0, ld, z2, g, c, true);
	packOperands<true>(z2, h);
	f32x4 y[4];
	layerOut(net->w3f, h, g, c, y);

	/* loss: mean over batch and channels of (y - t)^2 */
	float dy[3], l2 = 0.0f;
	const float scale = 2.0f / (3.0f * (float) A.batch);
#pragma unroll
	for (int ch = 0; ch < 3; ch++)
	{
		const float e = ownOutput(y, ch, lane) + net->b3[ch] - t[ch];
		l2 += e * e;
		dy[ch] = e * scale;
		A.dyt[(size_t) ch * ld + s] = (__bf16) dy[ch];
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		l2 += __shfl_xor(l2, off, 64);
	if (lane == 0)
		A.lossPart[blockIdx.x] = l2;

	/* dH2^T = W3^T dY^T: K = channels, lanes of g = 0 hold channels 0..2 */
	bf16x8 by[4];
#pragma unroll
	for (int nb = 0; nb < 4; nb++)
#pragma unroll
		for (int q = 0; q < 8; q++)
		{
			const float v = q < 3 ? __shfl(dy[q < 3 ? q : 0], 16 * nb + c, 64) : 0.0f;
			by[nb][q] = (__bf16) ((g == 0 && q < 3) ? v : 0.0f);
		}
	f32x4 dz[4][4];
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
	{
		const bf16x8 a = load8(net->w3b + (16 * mb + c) * 32 + 8 * g);
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
			dz[mb][nb] = mma(a, by[nb], zero4());
	}
	/* dZ2 = dH2 * relu'(Z2) */
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
#pragma unroll
			for (int r = 0; r < 4; r++)
				dz[mb][nb][r] = z2[mb][nb][r] > 0.0f ? dz[mb][nb][r] : 0.0f;
	storeTile(A.dz2t + s0, ld, dz, g, c, false);
	packOperands<false>(dz, h);
	/* dH1^T = W2^T dZ2^T; dZ1 = dH1 * relu'(Z1) */
	layerHidden(net->w2b, h, g, c, dz);
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
#pragma unroll
			for (int r = 0; r < 4; r++)
				dz[mb][nb][r] = z1[mb][nb][r] > 0.0f ? dz[mb][nb][r] : 0.0f;
	storeTile(A.dz1t + s0, ld, dz, g, c, false);
}

/* One weight-gradient tile: rows 16mb.. of P times rows 16(nb0+nb).. of Q
 * over one chunk of samples, written to partial[chunk][out + row*ld + col]. */
struct GradJob
{
	const __bf16* P;
	const __bf16* Q;
	uint32_t mb, nb0, nnb, rows, cols, out, ld;
};

__global__ __launch_bounds__(64) void k_weight_grads(const GradJob* __restrict__ jobs, uint32_t nJobs, uint32_t batch,
                                                     float* __restrict__ partial)
{
	const int lane = (int) threadIdx.x;
	const int g = lane >> 4, c = lane & 15;
	const GradJob J = jobs[blockIdx.x % nJobs];
	const uint32_t chunk = blockIdx.x / nJobs;
	const size_t k0 = (size_t) chunk * kChunk;
	f32x4 acc[4] = {zero4(), zero4(), zero4(), zero4()};
	const uint16_t* P = reinterpret_cast<const uint16_t*>(J.P) + (size_t) (16 * J.mb + c) * batch + k0 + 8 * g;
	const uint16_t* Q = reinterpret_cast<const uint16_t*>(J.Q) + (size_t) (16 * J.nb0 + c) * batch + k0 + 8 * g;
	for (int k = 0; k < kChunk; k += 32)
	{
		const bf16x8 a = load8(P + k);
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
			if (nb < (int) J.nnb)
				acc[nb] = mma(a, load8(Q + (size_t) 16 * nb * batch + k), acc[nb]);
	}
	float* dst = partial + (size_t) chunk * SRP_NN_PARAMS + J.out;
#pragma unroll
	for (int nb = 0; nb < 4; nb++)
		if (nb < (int) J.nnb)
#pragma unroll
			for (int r = 0; r < 4; r++)
			{
				const uint32_t row = 16 * J.mb + 4 * g + r, col = 16 * (J.nb0 + nb) + c;
				if (row < J.rows && col < J.cols)
					dst[row * J.ld + col] = acc[nb][r];
			}
}

struct NetMem
{
	float* w;      /* master, canonical */
	uint16_t* w1;
	uint16_t* w2f;
	uint16_t* w3f;
	uint16_t* w2b;
	uint16_t* w3b;
};

/* bf16 copies of canonical parameter p (biases are read from the master). */
__device__ inline void publish(const NetMem& M, int p, float v)
{
	const uint16_t b = __builtin_bit_cast(uint16_t, (__bf16) v);
	if (p < SRP_NN_OFF_B1)
		M.w1[p] = b;
	else if (p >= SRP_NN_OFF_W2 && p < SRP_NN_OFF_B2)
	{
		const int i = p - SRP_NN_OFF_W2, m = i / SRP_NN_HID, k = i % SRP_NN_HID;
		M.w2f[m * SRP_NN_HID + positionOf(k)] = b;
		M.w2b[k * SRP_NN_HID + positionOf(m)] = b;
	}
	else if (p >= SRP_NN_OFF_W3 && p < SRP_NN_OFF_B3)
	{
		const int i = p - SRP_NN_OFF_W3, ch = i / SRP_NN_HID, k = i % SRP_NN_HID;
		M.w3f[ch * SRP_NN_HID + positionOf(k)] = b;
		M.w3b[k * 32 + ch] = b;
	}
}

__global__ void k_publish(NetMem M)
{
	const int p = (int) (blockIdx.x * blockDim.x + threadIdx.x);
	if (p < SRP_NN_PARAMS)
		publish(M, p, M.w[p]);
}

struct AdamArgs
{
	NetMem M;
	float* m;
	float* v;
	const float* partial;
	uint32_t chunks;
	float lr, corr1, corr2;   /* corr = 1 / (1 - beta^t) */
	const float* lossPart;
	uint32_t nLoss;
	float lossScale;
	float* lossHist;          /* may be null */
	uint32_t step;
};

__global__ __launch_bounds__(256) void k_adam(AdamArgs A)
{
	const int p = (int) (blockIdx.x * blockDim.x + threadIdx.x);
	if (p < SRP_NN_PARAMS)
	{
		float g = 0.0f;
		for (uint32_t ch = 0; ch < A.chunks; ch++)
			g += A.partial[(size_t) ch * SRP_NN_PARAMS + p];
		const float m = 0.9f * A.m[p] + 0.1f * g;
		const float v = 0.999f * A.v[p] + 0.001f * g * g;
		A.m[p] = m;
		A.v[p] = v;
		const float w = A.M.w[p] - A.lr * (m * A.corr1) / (sqrtf(v * A.corr2) + 1e-8f);
		A.M.w[p] = w;
		publish(A.M, p, w);
	}
	if (blockIdx.x == 0 && A.lossHist && threadIdx.x < 64)
	{
		float s = 0.0f;
		for (uint32_t i = threadIdx.x; i < A.nLoss; i += 64)
			s += A.lossPart[i];
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
			s += __shfl_xor(s, off, 64);
		if (threadIdx.x == 0)
			A.lossHist[A.step] = s * A.lossScale;
	}
}

__global__ void k_sum_chunks(const float* __restrict__ partial, uint32_t chunks, float* __restrict__ grad)
{
	const int p = (int) (blockIdx.x * blockDim.x + threadIdx.x);
	if (p < SRP_NN_PARAMS)
	{
		float g = 0.0f;
		for (uint32_t ch = 0; ch < chunks; ch++)
			g += partial[(size_t) ch * SRP_NN_PARAMS + p];
		grad[p] = g;
	}
}

template <bool LANES>
__global__ __launch_bounds__(64) void k_infer(const SrpNnDevice* net, const float* __restrict__ feat, uint64_t n,
                                              float* __restrict__ rgb)
{
	const uint64_t s = (uint64_t) blockIdx.x * 64 + threadIdx.x;
	float xin[SRP_NN_IN];
#pragma unroll
	for (int d = 0; d < SRP_NN_IN; d++)
		xin[d] = s < n ? feat[s * SRP_NN_IN + d] : 0.0f;
	float o[3];
	if (LANES)
		inferLane(net, xin, o);
	else
		inferWave(net, xin, o);
	if (s < n)
#pragma unroll
		for (int ch = 0; ch < 3; ch++)
			rgb[s * 3 + ch] = o[ch];
}

/* -------------------------------------------------------- capture kernels */

__device__ inline bool keepPixel(uint64_t seed, size_t i, uint32_t threshold, const float* depth)
{
	if (!(depth[i] < 1.0f))
		return false;
	const uint32_t h = (uint32_t) (mix64(seed * 0x9E3779B97F4A7C15ull + i + 1) >> 32);
	return h < threshold || threshold == 0xFFFFFFFFu;
}

__global__ void k_capture_flags(const float* depth, size_t n, uint64_t seed, uint32_t threshold, uint32_t* flags)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flags[i] = keepPixel(seed, i, threshold, depth) ? 1u : 0u;
}

struct CaptureArgs
{
	const uint32_t* slot[5];
	const float* depth;
	const uint32_t* offsets;   /* exclusive scan of the flags */
	size_t n;
	uint64_t seed;
	uint32_t threshold;
	float* feat;
	float* targ;
	size_t base, capacity;
};

__global__ void k_capture_write(CaptureArgs A)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A.n || !keepPixel(A.seed, i, A.threshold, A.depth))
		return;
	const size_t o = A.base + A.offsets[i];
	if (o >= A.capacity)
		return;
	float f[8];
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const uint32_t px = A.slot[k][i];   /* R<<24 | G<<16 | B<<8 | A */
		f[2 * k] = (float) (px >> 16) / 65535.0f;
		f[2 * k + 1] = (float) (px & 0xFFFFu) / 65535.0f;
	}
#pragma unroll
	for (int k = 3; k < 6; k++)
		f[k] = f[k] * 2.0f - 1.0f;
#pragma unroll
	for (int k = 0; k < 8; k++)
		A.feat[o * 8 + k] = f[k];
	const uint32_t px = A.slot[4][i];
	A.targ[o * 3 + 0] = (float) (px >> 24) / 255.0f;
	A.targ[o * 3 + 1] = (float) ((px >> 16) & 255u) / 255.0f;
	A.targ[o * 3 + 2] = (float) ((px >> 8) & 255u) / 255.0f;
}

/* Sum of squared RGB differences, largest difference and pixel count, per
 * block: part[3*block + {0, 1, 2}].  The grid strides over the pixels and
 * the host adds the blocks' partials (no atomics on one shared address,
 * which serialised a grid of one block per 256 pixels). */
constexpr unsigned kCompareBlocks = 1024;

__global__ __launch_bounds__(256) void k_compare(const uint32_t* a, const uint32_t* b, const float* depth, size_t n,
                                                 unsigned long long* part)
{
	unsigned long long sse = 0, cnt = 0;
	uint32_t mx = 0;
	for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
	{
		if (!(depth[i] < 1.0f))
			continue;
		cnt++;
		const uint32_t pa = a[i], pb = b[i];
#pragma unroll
		for (int sh = 8; sh <= 24; sh += 8)
		{
			const int d = (int) ((pa >> sh) & 255u) - (int) ((pb >> sh) & 255u);
			sse += (unsigned long long) (d * d);
			const uint32_t ad = (uint32_t) (d < 0 ? -d : d);
			mx = ad > mx ? ad : mx;
		}
	}
	__shared__ unsigned long long sSse[4], sCnt[4];
	__shared__ uint32_t sMx[4];
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
	{
		sse += __shfl_xor(sse, off, 64);
		cnt += __shfl_xor(cnt, off, 64);
		const uint32_t o = __shfl_xor(mx, off, 64);
		mx = o > mx ? o : mx;
	}
	const int w = (int) (threadIdx.x >> 6);
	if ((threadIdx.x & 63u) == 0)
	{
		sSse[w] = sse;
		sCnt[w] = cnt;
		sMx[w] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int k = 1; k < 4; k++)
		{
			sse += sSse[k];
			cnt += sCnt[k];
			mx = sMx[k] > mx ? sMx[k] : mx;
		}
		part[3 * blockIdx.x + 0] = sse;
		part[3 * blockIdx.x + 1] = mx;
		part[3 * blockIdx.x + 2] = cnt;
	}
}

template <typename T>
T* devAlloc(size_t n)
{
	void* p = nullptr;
	NN_CHECK(hipMalloc(&p, n * sizeof(T) + 16));
	return (T*) p;
}

}  // namespace

/* Scratch of srpNeuralCompareNoAlloc: the per-block partial sums (squared
 * error, largest difference, pixel count) of at most kCompareBlocks blocks,
 * 24 KiB, which every GPU the library runs on holds until the process ends.
 * In host code this name is only a stand-in for the symbol: its device
 * address on each GPU comes from hipGetSymbolAddress (scratchAddr). */
static __device__ __align__(256) unsigned long long scratch[3 * kCompareBlocks];

namespace {

constexpr int kMaxScratchDevices = 64;
unsigned long long* g_scratchAddr[kMaxScratchDevices];     /* per device, set on first use */
unsigned long long g_scratchHost[3 * kCompareBlocks];      /* host copy of the partials */

/* The device address of `scratch` on the current device (looked up once per
 * device, with that device current); null, with a message, on failure. */
unsigned long long* scratchAddr(const char* fn)
{
	int dev = -1;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, fn, "No usable current device (%d)", dev);
		return nullptr;
	}
	if (!g_scratchAddr[dev])
	{
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
		{
			libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, fn, "Device scratch not available: %s",
			           hipGetErrorString(e));
			return nullptr;
		}
		g_scratchAddr[dev] = (unsigned long long*) p;
	}
	return g_scratchAddr[dev];
}

/* PSNR, largest difference and pixel count from the blocks' partials. */
void compareResult(const unsigned long long* hp, unsigned blocks, double* psnr, double* maxAbs, uint64_t* pixels)
{
	unsigned long long h[3] = {0, 0, 0};
	for (unsigned k = 0; k < blocks; k++)
	{
		h[0] += hp[3 * k + 0];
		h[1] = hp[3 * k + 1] > h[1] ? hp[3 * k + 1] : h[1];
		h[2] += hp[3 * k + 2];
	}
	const double mse = h[2] ? (double) h[0] / (3.0 * (double) h[2]) : 0.0;
	if (psnr) *psnr = mse > 0.0 ? 10.0 * log10(255.0 * 255.0 / mse) : INFINITY;
	if (maxAbs) *maxAbs = (double) h[1];
	if (pixels) *pixels = h[2];
}

}  // namespace

/* -------------------------------------------------------------- host objects */

struct SRPNeuralDataset
{
	float* feat = nullptr;
	float* targ = nullptr;
	size_t capacity = 0, size = 0;
	uint32_t* slot[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
	size_t slotPixels = 0;
	uint32_t* flags = nullptr;
	uint32_t* total = nullptr;
	srp::lib::ScanScratch scan;
};

struct SRPNeuralNet
{
	NetMem M;
	float* m = nullptr;
	float* v = nullptr;
	SrpNnDevice* view = nullptr;   /* device */
	SrpNnDevice host;              /* host copy of *view */
	int adamSteps = 0;
	/* training buffers, sized for kMaxBatch */
	__bf16 *xt = nullptr, *h1t = nullptr, *h2t = nullptr, *dz1t = nullptr, *dz2t = nullptr, *dyt = nullptr,
	       *ones = nullptr;
	float* partial = nullptr;
	float* lossPart = nullptr;
	float* lossHist = nullptr;
	size_t lossHistCap = 0;
	GradJob* jobs = nullptr;      /* device, rebuilt per batch size */
	uint32_t nJobs = 0, jobsBatch = 0;
	float* grad = nullptr;
	float* rgb = nullptr;
	size_t rgbCap = 0;
};

namespace {

bool checkBatch(const char* fn, size_t batch)
{
	if (batch == 0 || batch % kChunk != 0 || batch > (size_t) kMaxBatch)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, fn,
		           "Batch of %zu samples: must be a positive multiple of %d, at most %d", batch, kChunk, kMaxBatch);
		return false;
	}
	return true;
}

void buildJobs(SRPNeuralNet* N, uint32_t batch)
{
	if (N->jobsBatch == batch)
		return;
	std::vector<GradJob> J;
	auto add = [&](const __bf16* P, const __bf16* Q, uint32_t mb, uint32_t nb0, uint32_t nnb, uint32_t rows,
	               uint32_t cols, uint32_t out, uint32_t ld) {
		J.push_back(GradJob{P, Q, mb, nb0, nnb, rows, cols, out, ld});
	};
	for (uint32_t mb = 0; mb < 4; mb++)
		for (uint32_t grp = 0; grp < 8; grp++)
			add(N->dz1t, N->xt, mb, 4 * grp, 4, 64, 512, SRP_NN_OFF_W1, 512);
	for (uint32_t mb = 0; mb < 4; mb++)
		add(N->dz2t, N->h1t, mb, 0, 4, 64, 64, SRP_NN_OFF_W2, 64);
	add(N->dyt, N->h2t, 0, 0, 4, 3, 64, SRP_NN_OFF_W3, 64);
	for (uint32_t mb = 0; mb < 4; mb++)
		add(N->dz1t, N->ones, mb, 0, 1, 64, 1, SRP_NN_OFF_B1, 1);
	for (uint32_t mb = 0; mb < 4; mb++)
		add(N->dz2t, N->ones, mb, 0, 1, 64, 1, SRP_NN_OFF_B2, 1);
	add(N->dyt, N->ones, 0, 0, 1, 3, 1, SRP_NN_OFF_B3, 1);
	if (!N->jobs)
		N->jobs = devAlloc<GradJob>(64);
	NN_CHECK(hipMemcpy(N->jobs, J.data(), J.size() * sizeof(GradJob), hipMemcpyHostToDevice));
	N->nJobs = (uint32_t) J.size();
	N->jobsBatch = batch;
}

/* One forward/backward pass over `batch` samples and the chunked weight
 * gradients (partial sums in N->partial). */
void gradientPass(SRPNeuralNet* N, const SRPNeuralDataset* D, uint32_t batch, bool sequential, uint32_t step,
                  uint64_t seed, hipStream_t s)
{
	buildJobs(N, batch);
	TrainArgs A;
	A.feat = D->feat;
	A.targ = D->targ;
	A.n = D->size;
	A.sequential = sequential ? 1u : 0u;
	A.step = step;
	A.seed = seed;
	A.net = N->view;
	A.batch = batch;
	A.xt = N->xt; A.h1t = N->h1t; A.h2t = N->h2t; A.dz1t = N->dz1t; A.dz2t = N->dz2t; A.dyt = N->dyt;
	A.lossPart = N->lossPart;
	hipLaunchKernelGGL(k_train_pass, dim3(batch / 64), dim3(64), 0, s, A);
	hipLaunchKernelGGL(k_weight_grads, dim3(N->nJobs * (batch / kChunk)), dim3(64), 0, s, (const GradJob*) N->jobs,
	                   N->nJobs, batch, N->partial);
}

void uploadParams(SRPNeuralNet* N, const float* host)
{
	hipStream_t s = libStream();
	NN_CHECK(hipMemcpyAsync(N->M.w, host, SRP_NN_PARAMS * sizeof(float), hipMemcpyHostToDevice, s));
	NN_CHECK(hipMemsetAsync(N->m, 0, SRP_NN_PARAMS * sizeof(float), s));
	NN_CHECK(hipMemsetAsync(N->v, 0, SRP_NN_PARAMS * sizeof(float), s));
	hipLaunchKernelGGL(k_publish, dim3((SRP_NN_PARAMS + 255) / 256), dim3(256), 0, s, N->M);
	NN_CHECK(hipStreamSynchronize(s));
	N->adamSteps = 0;
}

}  // namespace

/* ---------------------------------------------------------------- dataset */

SRPNeuralDataset* srpNewNeuralDataset(size_t capacity)
{
	if (capacity == 0)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Capacity must be positive");
		return NULL;
	}
	SRPNeuralDataset* D = new SRPNeuralDataset();
	D->capacity = capacity;
	D->feat = devAlloc<float>(capacity * 8);
	D->targ = devAlloc<float>(capacity * 3);
	D->total = devAlloc<uint32_t>(4);
	return D;
}

void srpFreeNeuralDataset(SRPNeuralDataset* D)
{
	if (!D)
		return;
	NN_CHECK(hipStreamSynchronize(libStream()));
	NN_CHECK(hipFree(D->feat));
	NN_CHECK(hipFree(D->targ));
	for (int k = 0; k < 5; k++)
		if (D->slot[k]) NN_CHECK(hipFree(D->slot[k]));
	if (D->flags) NN_CHECK(hipFree(D->flags));
	NN_CHECK(hipFree(D->total));
	if (D->scan.buf) NN_CHECK(hipFree(D->scan.buf));
	if (D->scan.buf64) NN_CHECK(hipFree(D->scan.buf64));
	delete D;
}

void srpNeuralDatasetStage(SRPNeuralDataset* D, const SRPFramebuffer* fb, int slot)
{
	if (!D || !fb || slot < 0 || slot > 4)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid dataset, framebuffer or slot (%d)", slot);
		return;
	}
	if (D->slotPixels != fb->size)
	{
		NN_CHECK(hipStreamSynchronize(libStream()));
		for (int k = 0; k < 5; k++)
		{
			if (D->slot[k]) NN_CHECK(hipFree(D->slot[k]));
			D->slot[k] = devAlloc<uint32_t>(fb->size);
		}
		if (D->flags) NN_CHECK(hipFree(D->flags));
		D->flags = devAlloc<uint32_t>(fb->size);
		D->slotPixels = fb->size;
	}
	uint32_t* color;
	srp::lib::framebufferDevice(fb, &color, NULL, NULL);
	NN_CHECK(hipMemcpyAsync(D->slot[slot], color, fb->size * 4, hipMemcpyDeviceToDevice, libStream()));
}

size_t srpNeuralDatasetAppend(SRPNeuralDataset* D, const SRPFramebuffer* fb, float keep, uint64_t seed)
{
	if (!D || !fb || D->slotPixels != fb->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Nothing staged for a framebuffer of this size");
		return 0;
	}
	hipStream_t s = libStream();
	float* depth;
	srp::lib::framebufferDevice(fb, NULL, &depth, NULL);
	const double k = keep < 0.0f ? 0.0 : (keep > 1.0f ? 1.0 : (double) keep);
	const uint32_t threshold = k >= 1.0 ? 0xFFFFFFFFu : (uint32_t) (k * 4294967296.0);
	const size_t n = fb->size;
	const unsigned blocks = (unsigned) ((n + 255) / 256);
	hipLaunchKernelGGL(k_capture_flags, dim3(blocks), dim3(256), 0, s, (const float*) depth, n, seed, threshold, D->flags);
	srp::lib::scanExclusive(D->flags, n, D->total, D->scan, s);
	CaptureArgs A;
	for (int i = 0; i < 5; i++)
		A.slot[i] = D->slot[i];
	A.depth = depth;
	A.offsets = D->flags;
	A.n = n;
	A.seed = seed;
	A.threshold = threshold;
	A.feat = D->feat;
	A.targ = D->targ;
	A.base = D->size;
	A.capacity = D->capacity;
	hipLaunchKernelGGL(k_capture_write, dim3(blocks), dim3(256), 0, s, A);
	uint32_t total = 0;
	NN_CHECK(hipMemcpyAsync(&total, D->total, 4, hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
	const size_t added = (D->size + total > D->capacity) ? D->capacity - D->size : total;
	D->size += added;
	return added;
}

size_t srpNeuralDatasetSize(const SRPNeuralDataset* D) { return D ? D->size : 0; }

void srpNeuralDatasetRead(const SRPNeuralDataset* D, size_t first, size_t count, float* features, float* targets)
{
	if (!D || first + count > D->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Range out of bounds");
		return;
	}
	hipStream_t s = libStream();
	if (features)
		NN_CHECK(hipMemcpyAsync(features, D->feat + first * 8, count * 8 * sizeof(float), hipMemcpyDeviceToHost, s));
	if (targets)
		NN_CHECK(hipMemcpyAsync(targets, D->targ + first * 3, count * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
}

void srpNeuralDatasetWrite(SRPNeuralDataset* D, const float* features, const float* targets, size_t count)
{
	if (!D || count > D->capacity)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "%zu samples exceed the capacity", count);
		return;
	}
	hipStream_t s = libStream();
	NN_CHECK(hipMemcpyAsync(D->feat, features, count * 8 * sizeof(float), hipMemcpyHostToDevice, s));
	NN_CHECK(hipMemcpyAsync(D->targ, targets, count * 3 * sizeof(float), hipMemcpyHostToDevice, s));
	NN_CHECK(hipStreamSynchronize(s));
	D->size = count;
}

/* ---------------------------------------------------------------- network */

SRPNeuralNet* srpNewNeuralNet(uint64_t seed)
{
	SRPNeuralNet* N = new SRPNeuralNet();
	N->M.w = devAlloc<float>(SRP_NN_PARAMS);
	N->m = devAlloc<float>(SRP_NN_PARAMS);
	N->v = devAlloc<float>(SRP_NN_PARAMS);
	N->M.w1 = devAlloc<uint16_t>(64 * 512);
	N->M.w2f = devAlloc<uint16_t>(64 * 64);
	N->M.w3f = devAlloc<uint16_t>(16 * 64);
	N->M.w2b = devAlloc<uint16_t>(64 * 64);
	N->M.w3b = devAlloc<uint16_t>(64 * 32);
	NN_CHECK(hipMemset(N->M.w3f, 0, 16 * 64 * 2));
	NN_CHECK(hipMemset(N->M.w3b, 0, 64 * 32 * 2));
	const size_t B = kMaxBatch;
	N->xt = devAlloc<__bf16>(512 * B);
	N->h1t = devAlloc<__bf16>(64 * B);
	N->h2t = devAlloc<__bf16>(64 * B);
	N->dz1t = devAlloc<__bf16>(64 * B);
	N->dz2t = devAlloc<__bf16>(64 * B);
	N->dyt = devAlloc<__bf16>(16 * B);
	N->ones = devAlloc<__bf16>(16 * B);
	NN_CHECK(hipMemset(N->dyt, 0, 16 * B * 2));
	N->partial = devAlloc<float>((size_t) kMaxChunks * SRP_NN_PARAMS);
	N->lossPart = devAlloc<float>(B / 64);
	N->grad = devAlloc<float>(SRP_NN_PARAMS);
	{
		std::vector<uint16_t> ones(16 * B, 0);
		for (size_t i = 0; i < B; i++)
			ones[i] = 0x3F80;  /* bf16 1.0 in row 0 */
		NN_CHECK(hipMemcpy(N->ones, ones.data(), ones.size() * 2, hipMemcpyHostToDevice));
	}
	SrpNnDevice view;
	view.w1 = N->M.w1; view.w2f = N->M.w2f; view.w3f = N->M.w3f; view.w2b = N->M.w2b; view.w3b = N->M.w3b;
	view.b1 = N->M.w + SRP_NN_OFF_B1;
	view.b2 = N->M.w + SRP_NN_OFF_B2;
	view.b3 = N->M.w + SRP_NN_OFF_B3;
	for (int d = 0; d < SRP_NN_IN; d++)
		view.step[d] = SRP_NN_FREQ_STEP;
	N->host = view;
	N->view = devAlloc<SrpNnDevice>(1);
	NN_CHECK(hipMemcpy(N->view, &view, sizeof(view), hipMemcpyHostToDevice));

	/* PyTorch-style default: weights and biases uniform in +-1/sqrt(fan_in). */
	std::vector<float> p(SRP_NN_PARAMS);
	uint64_t st = seed;
	auto uni = [&](float bound) {
		st += 0x9E3779B97F4A7C15ull;
		const double u = (double) (mix64(st) >> 11) * (1.0 / 9007199254740992.0);
		return (float) ((2.0 * u - 1.0) * bound);
	};
	const float b1 = 1.0f / sqrtf((float) SRP_NN_ENC), b2 = 1.0f / sqrtf((float) SRP_NN_HID);
	for (int i = 0; i < SRP_NN_PARAMS; i++)
		p[i] = uni(i < SRP_NN_OFF_W2 ? b1 : b2);
	uploadParams(N, p.data());
	return N;
}

void srpFreeNeuralNet(SRPNeuralNet* N)
{
	if (!N)
		return;
	NN_CHECK(hipStreamSynchronize(libStream()));
	void* bufs[] = {N->M.w, N->m, N->v, N->M.w1, N->M.w2f, N->M.w3f, N->M.w2b, N->M.w3b, N->xt, N->h1t, N->h2t,
	                N->dz1t, N->dz2t, N->dyt, N->ones, N->partial, N->lossPart, N->lossHist, N->jobs, N->grad,
	                N->rgb, N->view};
	for (void* b : bufs)
		if (b) NN_CHECK(hipFree(b));
	delete N;
}

void srpNeuralNetGetParams(const SRPNeuralNet* N, float* params)
{
	if (!N || !params)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid network or parameter array");
		return;
	}
	hipStream_t s = libStream();
	NN_CHECK(hipMemcpyAsync(params, N->M.w, SRP_NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
}

void srpNeuralNetSetParams(SRPNeuralNet* N, const float* params)
{
	if (!N || !params)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid network or parameter array");
		return;
	}
	uploadParams(N, params);
}

const void* srpNeuralNetDevice(const SRPNeuralNet* N) { return N ? (const void*) N->view : NULL; }

void srpNeuralNetSetEncoding(SRPNeuralNet* N, const float* step)
{
	if (!step)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid step array");
		return;
	}
	for (int d = 0; d < SRP_NN_IN; d++)
		if (!(step[d] > 0.0f && step[d] < 1024.0f))
		{
			libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__,
			           "frequency step %g of feature %d is not in (0, 1024)", (double) step[d], d);
			return;
		}
	if (!N)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid network");
		return;
	}
	hipStream_t s = libStream();
	NN_CHECK(hipStreamSynchronize(s));
	for (int d = 0; d < SRP_NN_IN; d++)
		N->host.step[d] = step[d];
	NN_CHECK(hipMemcpy(N->view, &N->host, sizeof(SrpNnDevice), hipMemcpyHostToDevice));
}

void srpNeuralNetGetEncoding(const SRPNeuralNet* N, float* step)
{
	if (!N || !step)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid network or step array");
		return;
	}
	for (int d = 0; d < SRP_NN_IN; d++)
		step[d] = N->host.step[d];
}

void srpNeuralNetTrain(SRPNeuralNet* N, const SRPNeuralDataset* D, int steps, int batch, float lr0, float lr1,
                       uint64_t seed, float* lossOut)
{
	if (!N || !D || D->size == 0 || steps <= 0)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Nothing to train on");
		return;
	}
	if (!checkBatch(__func__, (size_t) batch))
		return;
	hipStream_t s = libStream();
	if (lossOut && N->lossHistCap < (size_t) steps)
	{
		NN_CHECK(hipStreamSynchronize(s));
		if (N->lossHist) NN_CHECK(hipFree(N->lossHist));
		N->lossHist = devAlloc<float>((size_t) steps);
		N->lossHistCap = (size_t) steps;
	}
	for (int t = 0; t < steps; t++)
	{
		gradientPass(N, D, (uint32_t) batch, false, (uint32_t) t, seed, s);
		N->adamSteps++;
		const double cosw = 0.5 * (1.0 + cos(M_PI * (double) t / (double) steps));
		AdamArgs A;
		A.M = N->M;
		A.m = N->m;
		A.v = N->v;
		A.partial = N->partial;
		A.chunks = (uint32_t) (batch / kChunk);
		A.lr = (float) (lr1 + (lr0 - lr1) * cosw);
		A.corr1 = (float) (1.0 / (1.0 - pow(0.9, N->adamSteps)));
		A.corr2 = (float) (1.0 / (1.0 - pow(0.999, N->adamSteps)));
		A.lossPart = N->lossPart;
		A.nLoss = (uint32_t) (batch / 64);
		A.lossScale = 1.0f / (3.0f * (float) batch);
		A.lossHist = lossOut ? N->lossHist : nullptr;
		A.step = (uint32_t) t;
		hipLaunchKernelGGL(k_adam, dim3((SRP_NN_PARAMS + 255) / 256), dim3(256), 0, s, A);
	}
	if (lossOut)
		NN_CHECK(hipMemcpyAsync(lossOut, N->lossHist, (size_t) steps * sizeof(float), hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
}

void srpNeuralNetGradients(SRPNeuralNet* N, const SRPNeuralDataset* D, size_t count, float* grad, float* loss)
{
	if (!N || !D || count > D->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Count exceeds the dataset");
		return;
	}
	if (!checkBatch(__func__, count))
		return;
	hipStream_t s = libStream();
	gradientPass(N, D, (uint32_t) count, true, 0, 0, s);
	hipLaunchKernelGGL(k_sum_chunks, dim3((SRP_NN_PARAMS + 255) / 256), dim3(256), 0, s, (const float*) N->partial,
	                   (uint32_t) (count / kChunk), N->grad);
	NN_CHECK(hipMemcpyAsync(grad, N->grad, SRP_NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s));
	std::vector<float> lp(count / 64);
	NN_CHECK(hipMemcpyAsync(lp.data(), N->lossPart, lp.size() * sizeof(float), hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
	if (loss)
	{
		double sum = 0.0;
		for (float v : lp)
			sum += v;
		*loss = (float) (sum / (3.0 * (double) count));
	}
}

void srpNeuralNetInfer(const SRPNeuralNet* Nc, const SRPNeuralDataset* D, size_t count, float* rgb, int lanes)
{
	SRPNeuralNet* N = const_cast<SRPNeuralNet*>(Nc);
	if (!N || !D || count > D->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Count exceeds the dataset");
		return;
	}
	if (count == 0)
		return;
	hipStream_t s = libStream();
	if (N->rgbCap < count)
	{
		NN_CHECK(hipStreamSynchronize(s));
		if (N->rgb) NN_CHECK(hipFree(N->rgb));
		N->rgb = devAlloc<float>(count * 3);
		N->rgbCap = count;
	}
	const unsigned blocks = (unsigned) ((count + 63) / 64);
	if (lanes)
		hipLaunchKernelGGL(k_infer<true>, dim3(blocks), dim3(64), 0, s, (const SrpNnDevice*) N->view,
		                   (const float*) D->feat, (uint64_t) count, N->rgb);
	else
		hipLaunchKernelGGL(k_infer<false>, dim3(blocks), dim3(64), 0, s, (const SrpNnDevice*) N->view,
		                   (const float*) D->feat, (uint64_t) count, N->rgb);
	NN_CHECK(hipMemcpyAsync(rgb, N->rgb, count * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
}

/* Device time of the matrix-core inference over dataset samples [0, count):
 * mean milliseconds per pass over `reps` passes (HIP events on the library
 * stream around the kernels only; the first pass is a warm-up). */
double srpNeuralNetTimeInference(const SRPNeuralNet* Nc, const SRPNeuralDataset* D, size_t count, int reps)
{
	SRPNeuralNet* N = const_cast<SRPNeuralNet*>(Nc);
	if (!N || !D || count == 0 || count > D->size || reps < 1)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid network, dataset, count or repetitions");
		return 0.0;
	}
	hipStream_t s = libStream();
	if (N->rgbCap < count)
	{
		NN_CHECK(hipStreamSynchronize(s));
		if (N->rgb) NN_CHECK(hipFree(N->rgb));
		N->rgb = devAlloc<float>(count * 3);
		N->rgbCap = count;
	}
	const unsigned blocks = (unsigned) ((count + 63) / 64);
	hipEvent_t a, b;
	NN_CHECK(hipEventCreate(&a));
	NN_CHECK(hipEventCreate(&b));
	hipLaunchKernelGGL(k_infer<false>, dim3(blocks), dim3(64), 0, s, (const SrpNnDevice*) N->view,
	                   (const float*) D->feat, (uint64_t) count, N->rgb);
	NN_CHECK(hipEventRecord(a, s));
	for (int r = 0; r < reps; r++)
		hipLaunchKernelGGL(k_infer<false>, dim3(blocks), dim3(64), 0, s, (const SrpNnDevice*) N->view,
		                   (const float*) D->feat, (uint64_t) count, N->rgb);
	NN_CHECK(hipEventRecord(b, s));
	NN_CHECK(hipEventSynchronize(b));
	float ms = 0.0f;
	NN_CHECK(hipEventElapsedTime(&ms, a, b));
	NN_CHECK(hipEventDestroy(a));
	NN_CHECK(hipEventDestroy(b));
	return (double) ms / reps;
}

/* ------------------------------------------------------------- evaluation */

void srpNeuralCompare(const SRPFramebuffer* a, const SRPFramebuffer* b, const SRPFramebuffer* mask, double* psnr,
                      double* maxAbs, uint64_t* pixels)
{
	if (!a || !b || !mask || a->size != b->size || a->size != mask->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Framebuffers must exist and have one size");
		return;
	}
	hipStream_t s = libStream();
	uint32_t *ca, *cb;
	float* depth;
	srp::lib::framebufferDevice(a, &ca, NULL, NULL);
	srp::lib::framebufferDevice(b, &cb, NULL, NULL);
	srp::lib::framebufferDevice(mask, NULL, &depth, NULL);
	const size_t need = (a->size + 255) / 256;
	const unsigned blocks = need == 0 ? 1u : (need < kCompareBlocks ? (unsigned) need : kCompareBlocks);
	unsigned long long* part = devAlloc<unsigned long long>(3 * (size_t) blocks);
	hipLaunchKernelGGL(k_compare, dim3(blocks), dim3(256), 0, s, (const uint32_t*) ca, (const uint32_t*) cb,
	                   (const float*) depth, a->size, part);
	std::vector<unsigned long long> hp(3 * (size_t) blocks);
	NN_CHECK(hipMemcpyAsync(hp.data(), part, 3 * sizeof(unsigned long long) * blocks, hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
	NN_CHECK(hipFree(part));
	compareResult(hp.data(), blocks, psnr, maxAbs, pixels);
}

/* srpNeuralCompare without any allocation: the partials go to the file's
 * static device scratch (one per GPU) and a static host array.  Calls that
 * overlap on one GPU share both and race. */
void srpNeuralCompareNoAlloc(const SRPFramebuffer* a, const SRPFramebuffer* b, const SRPFramebuffer* mask,
                             double* psnr, double* maxAbs, uint64_t* pixels)
{
	if (!a || !b || !mask || a->size != b->size || a->size != mask->size)
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Framebuffers must exist and have one size");
		return;
	}
	hipStream_t s = libStream();
	unsigned long long* part = scratchAddr(__func__);
	if (!part)
		return;
	const size_t need = (a->size + 255) / 256;
	const unsigned blocks = need == 0 ? 1u : (need < kCompareBlocks ? (unsigned) need : kCompareBlocks);
	/* the grid strides over any frame, so its partials never exceed the
	 * scratch; a larger grid would be refused here */
	if (3 * (size_t) blocks > sizeof(g_scratchHost) / sizeof(g_scratchHost[0]))
	{
		libMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__,
		           "%u blocks of partials exceed the %zu-entry scratch", blocks,
		           sizeof(g_scratchHost) / sizeof(g_scratchHost[0]) / 3);
		return;
	}
	uint32_t *ca, *cb;
	float* depth;
	srp::lib::framebufferDevice(a, &ca, NULL, NULL);
	srp::lib::framebufferDevice(b, &cb, NULL, NULL);
	srp::lib::framebufferDevice(mask, NULL, &depth, NULL);
	/* every block writes its own three partials: nothing to clear */
	hipLaunchKernelGGL(k_compare, dim3(blocks), dim3(256), 0, s, (const uint32_t*) ca, (const uint32_t*) cb,
	                   (const float*) depth, a->size, part);
	NN_CHECK(hipMemcpyAsync(g_scratchHost, part, 3 * sizeof(unsigned long long) * blocks, hipMemcpyDeviceToHost, s));
	NN_CHECK(hipStreamSynchronize(s));
	compareResult(g_scratchHost, blocks, psnr, maxAbs, pixels);
}
