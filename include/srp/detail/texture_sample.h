/* Host+device texture lookup for HIP shader modules (included by
 * srp/texture.h under hipcc).  An SRPTexture* is the device address of an
 * SrpTexDesc; device code reads it directly, host code asks the library.
 *
 * Same arithmetic as the reference's core/texture.c:56-81: wrap in double,
 * texel coordinates in float, nearest texel by adding 0.5 and truncating,
 * 1/255 scaling in float, alpha 1.  Coordinates are clamped to the texture
 * (the reference reads out of bounds for NaN coordinates). */
#pragma once

#include <math.h>
#include "srp/device/abi.h"

SRP_EXTERN_C_BEGIN
void srpTextureGetFilteredColorHost(const SRPTexture* texture, float u, float v, float out[4]);
SRP_EXTERN_C_END

static inline __host__ __device__ void srpTextureGetFilteredColor(const SRPTexture* tex, float u, float v, float out[4])
{
#if defined(__HIP_DEVICE_COMPILE__)
	const SrpTexDesc* d = (const SrpTexDesc*) tex;
	if (u < 0 || u > 1)
		u = (d->wrapX == TW_REPEAT) ? (float) ((double) u - floor((double) u)) : (float) fmax(0.0, fmin(1.0, (double) u));
	if (v < 0 || v > 1)
		v = (d->wrapY == TW_REPEAT) ? (float) ((double) v - floor((double) v)) : (float) fmax(0.0, fmin(1.0, (double) v));
	const float x = (float) d->wM1 * u;
	const float y = (float) d->hM1 * (1 - v);
	const double xd = (double) x + 0.5, yd = (double) y + 0.5;
	int64_t xi = (xd >= 0.0 && xd < 9.0e18) ? (int64_t) xd : 0;
	int64_t yi = (yd >= 0.0 && yd < 9.0e18) ? (int64_t) yd : 0;
	if (xi > d->wM1) xi = d->wM1;
	if (yi > d->hM1) yi = d->hM1;
	const uint32_t texel = d->texels[yi * (int64_t) d->w + xi];
	const float inv255 = (float) (1. / 255.);
	out[0] = (float) (int) (texel & 255u) * inv255;
	out[1] = (float) (int) ((texel >> 8) & 255u) * inv255;
	out[2] = (float) (int) ((texel >> 16) & 255u) * inv255;
	out[3] = 1.0f;
#else
	srpTextureGetFilteredColorHost(tex, u, v, out);
#endif
}
