/* Device side of the neural fragment shader: the Fourier encoding and the
 * 8 -> 512 -> 64 -> 64 -> 3 network on the matrix cores (bf16 operands,
 * fp32 accumulation), for one wave of 64 fragments at a time, plus a
 * per-lane fallback for kernels whose lanes are not convergent.
 *
 * Layouts: srp/neural_abi.h.  Matrix operands of the 16x16x32 bf16
 * product, for lane l with g = l/16 and c = l%16:
 *     a[q] = A[row c][k 8g+q],   b[q] = B[k 8g+q][col c],   d[r] = D[row 4g+r][col c].
 * Every layer is computed transposed (rows = units, columns = fragments),
 * so a layer's result tile, after bias and ReLU, is already the B operand
 * of the next layer once its K order is the permuted one of neural_abi.h. */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srp/neural_abi.h"

namespace srp {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p)
{
	return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ f32x4 zero4()
{
	f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
	return z;
}

/* Hidden unit at K position (ks, g, q), and its inverse. */
__host__ __device__ constexpr int hiddenAt(int ks, int g, int q)
{
	return 16 * (2 * ks + (q >> 2)) + 4 * g + (q & 3);
}
__host__ __device__ constexpr int positionOf(int h)
{
	return 32 * (h >> 5) + 8 * ((h >> 2) & 3) + 4 * ((h >> 4) & 1) + (h & 3);
}

/* sin and cos of 2*pi*r, r in revolutions (the hardware's native units). */
__device__ __forceinline__ void sincosRev(float r, float& s, float& c)
{
	r = r - floorf(r);
	s = __builtin_amdgcn_sinf(r);
	c = __builtin_amdgcn_cosf(r);
}

/* Encoding values of a feature at frequencies j0..j0+3, interleaved sin,
 * cos; u = the feature times its frequency step, (sd, cd) = sin, cos of
 * 2*pi*u.  The first pair is evaluated directly, the next three by
 * rotation. */
__device__ __forceinline__ void encode4(float u, int j0, float sd, float cd, float out[8])
{
	float s, c;
	sincosRev((float) (j0 + 1) * u, s, c);
	out[0] = s;
	out[1] = c;
#pragma unroll
	for (int t = 1; t < 4; t++)
	{
		const float s2 = __builtin_fmaf(s, cd, c * sd);
		const float c2 = __builtin_fmaf(c, cd, -(s * sd));
		s = s2;
		c = c2;
		out[2 * t] = s;
		out[2 * t + 1] = c;
	}
}

__device__ __forceinline__ bf16x8 encode8(float u, int j0, float sd, float cd)
{
	float v[8];
	encode4(u, j0, sd, cd, v);
	bf16x8 o;
#pragma unroll
	for (int q = 0; q < 8; q++)
		o[q] = (__bf16) v[q];
	return o;
}

/* Bias + ReLU of a 64 x 64 result tile acc[mb][nb] (rows 16mb+4g+r), packed
 * as the two K steps of the next layer's B operand. */
__device__ __forceinline__ void hiddenOperands(const f32x4 acc[4][4], const float* bias, int g, bf16x8 h[2][4])
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const float b = bias[16 * mb + 4 * g + r];
#pragma unroll
			for (int nb = 0; nb < 4; nb++)
			{
				const float v = acc[mb][nb][r] + b;
				h[mb >> 1][nb][4 * (mb & 1) + r] = (__bf16) (v > 0.0f ? v : 0.0f);
			}
		}
}

/* In-place bias of a 64 x 64 result tile, and its ReLU packed as the next
 * layer's B operand (pack without ReLU for gradients). */
__device__ __forceinline__ void biasInPlace(f32x4 acc[4][4], const float* bias, int g)
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const float b = bias[16 * mb + 4 * g + r];
#pragma unroll
			for (int nb = 0; nb < 4; nb++)
				acc[mb][nb][r] += b;
		}
}

template <bool RELU>
__device__ __forceinline__ void packOperands(const f32x4 acc[4][4], bf16x8 h[2][4])
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int r = 0; r < 4; r++)
#pragma unroll
			for (int nb = 0; nb < 4; nb++)
			{
				const float v = acc[mb][nb][r];
				h[mb >> 1][nb][4 * (mb & 1) + r] = (__bf16) (RELU ? (v > 0.0f ? v : 0.0f) : v);
			}
}

/* Layer 1 for the wave's 64 fragments: acc[mb][nb] = W1 X^T (no bias).
 * xin = this lane's own fragment's features, step = the network's
 * frequency steps; column 16nb + c of X^T is lane 16nb + c's fragment,
 * fetched by shuffle one feature at a time (so no lane holds all four
 * columns' features at once).  If xt != nullptr the encoded operand is also
 * stored transposed (xt[k * ld + col0 + 16nb + c]). */
__device__ __forceinline__ void layer1(const uint16_t* __restrict__ w1, const float* step, const float xin[SRP_NN_IN],
                                       int g, int c, f32x4 acc[4][4], __bf16* xt = nullptr, size_t ld = 0)
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
			acc[mb][nb] = zero4();
#pragma unroll 1
	for (int d = 0; d < SRP_NN_IN; d++)
	{
		/* xin[d] by selects (a dynamic index would put xin in memory) */
		float own = xin[0];
#pragma unroll
		for (int t = 1; t < SRP_NN_IN; t++)
			own = (d == t) ? xin[t] : own;
		const float sdim = step[d];
		float ud[4], sd[4], cd[4];
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
		{
			ud[nb] = sdim * __shfl(own, 16 * nb + c, 64);
			sincosRev(ud[nb], sd[nb], cd[nb]);
		}
#pragma unroll
		for (int half = 0; half < 2; half++)
		{
			const int ks = 2 * d + half;
			const int j0 = 16 * half + 4 * g;
			bf16x8 b[4];
#pragma unroll
			for (int nb = 0; nb < 4; nb++)
			{
				b[nb] = encode8(ud[nb], j0, sd[nb], cd[nb]);
				if (xt)
#pragma unroll
					for (int q = 0; q < 8; q++)
						xt[(size_t) (32 * ks + 8 * g + q) * ld + 16 * nb + c] = b[nb][q];
			}
#pragma unroll
			for (int mb = 0; mb < 4; mb++)
			{
				const bf16x8 a = load8(w1 + (16 * mb + c) * SRP_NN_ENC + 32 * ks + 8 * g);
#pragma unroll
				for (int nb = 0; nb < 4; nb++)
					acc[mb][nb] = mma(a, b[nb], acc[mb][nb]);
			}
		}
	}
}

/* acc[mb][nb] = A h, A = 64 x 64 permuted-K matrix (w2f or w2b). */
__device__ __forceinline__ void layerHidden(const uint16_t* __restrict__ A, const bf16x8 h[2][4], int g, int c,
                                            f32x4 acc[4][4])
{
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
			acc[mb][nb] = zero4();
#pragma unroll
	for (int ks = 0; ks < 2; ks++)
#pragma unroll
		for (int mb = 0; mb < 4; mb++)
		{
			const bf16x8 a = load8(A + (16 * mb + c) * SRP_NN_HID + 32 * ks + 8 * g);
#pragma unroll
			for (int nb = 0; nb < 4; nb++)
				acc[mb][nb] = mma(a, h[ks][nb], acc[mb][nb]);
		}
}

/* Output layer: y[nb][r] = (W3 h)[channel 4g+r][fragment 16nb+c] (no bias). */
__device__ __forceinline__ void layerOut(const uint16_t* __restrict__ w3f, const bf16x8 h[2][4], int g, int c, f32x4 y[4])
{
#pragma unroll
	for (int nb = 0; nb < 4; nb++)
		y[nb] = zero4();
#pragma unroll
	for (int ks = 0; ks < 2; ks++)
	{
		const bf16x8 a = load8(w3f + c * SRP_NN_HID + 32 * ks + 8 * g);
#pragma unroll
		for (int nb = 0; nb < 4; nb++)
			y[nb] = mma(a, h[ks][nb], y[nb]);
	}
}

/* Channel ch of this lane's own fragment from the output tile. */
__device__ __forceinline__ float ownOutput(const f32x4 y[4], int ch, int lane)
{
	float v = 0.0f;
#pragma unroll
	for (int nb = 0; nb < 4; nb++)
	{
		const float t = __shfl(y[nb][ch], lane & 15, 64);
		if (nb == (lane >> 4))
			v = t;
	}
	return v;
}

/* The network for the wave's 64 fragments.  All 64 lanes must call it
 * together; each passes its own fragment's features (any finite values on
 * lanes without one) and receives that fragment's RGB. */
__device__ inline void inferWave(const SrpNnDevice* __restrict__ net, const float xin[SRP_NN_IN], float rgb[3])
{
	const int lane = (int) (threadIdx.x & 63u);
	const int g = lane >> 4, c = lane & 15;
	f32x4 acc[4][4];
	layer1(net->w1, net->step, xin, g, c, acc);
	bf16x8 h[2][4];
	hiddenOperands(acc, net->b1, g, h);
	layerHidden(net->w2f, h, g, c, acc);
	hiddenOperands(acc, net->b2, g, h);
	f32x4 y[4];
	layerOut(net->w3f, h, g, c, y);
#pragma unroll
	for (int ch = 0; ch < 3; ch++)
		rgb[ch] = ownOutput(y, ch, lane) + net->b3[ch];
}

__device__ __forceinline__ float bf(uint16_t bits)
{
	return __uint_as_float((uint32_t) bits << 16);
}

__device__ __forceinline__ float roundBf(float v)
{
	return (float) (__bf16) v;
}

/* The same network for one fragment on one lane (vector ALU only): the
 * same bf16 operands, fp32 sums in sequential order. */
__device__ inline void inferLane(const SrpNnDevice* __restrict__ net, const float xin[SRP_NN_IN], float rgb[3])
{
	float h1[SRP_NN_HID];
#pragma unroll
	for (int m = 0; m < SRP_NN_HID; m++)
		h1[m] = 0.0f;
	for (int d = 0; d < SRP_NN_IN; d++)
	{
		const float u = net->step[d] * xin[d];
		float sd, cd;
		sincosRev(u, sd, cd);
		for (int j0 = 0; j0 < SRP_NN_FREQ; j0 += 4)
		{
			float e[8];
			encode4(u, j0, sd, cd, e);
			const int k0 = 64 * d + 2 * j0;
#pragma unroll
			for (int q = 0; q < 8; q++)
				e[q] = roundBf(e[q]);
			for (int m = 0; m < SRP_NN_HID; m++)
			{
				const uint16_t* w = net->w1 + m * SRP_NN_ENC + k0;
				float s = h1[m];
#pragma unroll
				for (int q = 0; q < 8; q++)
					s = __builtin_fmaf(bf(w[q]), e[q], s);
				h1[m] = s;
			}
		}
	}
	for (int m = 0; m < SRP_NN_HID; m++)
	{
		const float v = h1[m] + net->b1[m];
		h1[m] = roundBf(v > 0.0f ? v : 0.0f);
	}
	float h2[SRP_NN_HID];
	for (int m = 0; m < SRP_NN_HID; m++)
	{
		float s = 0.0f;
		for (int k = 0; k < SRP_NN_HID; k++)
			s = __builtin_fmaf(bf(net->w2f[m * SRP_NN_HID + positionOf(k)]), h1[k], s);
		const float v = s + net->b2[m];
		h2[m] = roundBf(v > 0.0f ? v : 0.0f);
	}
	for (int ch = 0; ch < 3; ch++)
	{
		float s = 0.0f;
		for (int k = 0; k < SRP_NN_HID; k++)
			s = __builtin_fmaf(bf(net->w3f[ch * SRP_NN_HID + positionOf(k)]), h2[k], s);
		rgb[ch] = s + net->b3[ch];
	}
}

}  // namespace nn
}  // namespace srp
