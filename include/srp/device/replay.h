/* The reference renderer walks a triangle's barycentrics by repeated float
 * addition (raster/triangle.c:92-107 of kitrofimov/srp): the value at a
 * pixel n steps from the bounding box origin is a + d + d + ... (n
 * additions, each rounded).  srpReplay returns exactly that sum without
 * performing the n additions one by one.
 *
 * Within one binade [2^E, 2^(E+1)) of one sign every float is an integer
 * multiple M of the binade's spacing u, and a + d rounds to the multiple
 * nearest to M + d/u.  While the exact sums stay inside the binade the sum
 * therefore moves by the same number of spacings at every step - except
 * when d/u lies exactly halfway between integers, where round-half-even
 * picks by parity; after one such step M has the parity that makes every
 * later step equal.  So: two real additions settle the step D (in units
 * of u), then as many further steps as keep M + k*D safely inside the
 * binade are applied at once, and the walk continues from there.  The
 * number of iterations is bounded by the binades the walk crosses; short
 * walks are summed directly.
 *
 * Host and device code share this header; FP contraction must be off. */
#pragma once

#include <stdint.h>
#include <string.h>
#include "srp/detail/qualifiers.h"

SRP_HD static inline uint32_t srpFloatBits_(float x)
{
	uint32_t b;
	memcpy(&b, &x, 4);
	return b;
}

SRP_HD static inline float srpBitsFloat_(uint32_t b)
{
	float x;
	memcpy(&x, &b, 4);
	return x;
}

/* Walks of at most this many steps are summed directly. */
#define SRP_REPLAY_DIRECT 24

/* a + d added n times in sequence (n >= 0), bit for bit. */
SRP_HD static inline float srpReplay(float a, float d, int n)
{
	while (n > SRP_REPLAY_DIRECT)
	{
		const uint32_t ab = srpFloatBits_(a);
		const uint32_t ex = (ab >> 23) & 0xFFu;
		if (d == 0.0f && ex != 0xFFu)
			return a + d;   /* one addition fixes the sign of a zero sum; later ones change nothing */
		const float b = a + d;
		const float c = b + d;
		n -= 2;
		/* zero, tiny, subnormal, infinite or NaN sums, or a binade change
		 * within the two steps: continue from c */
		const uint32_t bb = srpFloatBits_(b), cb = srpFloatBits_(c);
		const uint32_t sign = ab & 0x80000000u;
		if (ex < 24u || ex == 0xFFu || ((bb >> 23) & 0xFFu) != ex || ((cb >> 23) & 0xFFu) != ex ||
		    (bb & 0x80000000u) != sign || (cb & 0x80000000u) != sign)
		{
			a = c;
			continue;
		}
		/* magnitudes in units of the binade spacing (below 2^24) */
		const uint32_t mb = (bb & 0x7FFFFFu) | 0x800000u;
		const uint32_t mc = (cb & 0x7FFFFFu) | 0x800000u;
		if (mb == mc)
			return c;   /* c + d rounds back to c: the sum stays */
		/* steps that keep the magnitude within [2^23 + 1, 2^24 - 1] */
		uint32_t k;
		uint32_t m;
		if (mc > mb)
		{
			const uint32_t D = mc - mb;
			k = (0xFFFFFFu - mc) / D;
			if (k > (uint32_t) n) k = (uint32_t) n;
			m = mc + k * D;
		}
		else
		{
			const uint32_t D = mb - mc;
			k = mc > 0x800001u ? (mc - 0x800001u) / D : 0u;
			if (k > (uint32_t) n) k = (uint32_t) n;
			m = mc - k * D;
		}
		a = srpBitsFloat_(sign | (ex << 23) | (m & 0x7FFFFFu));
		n -= (int) k;
	}
	for (int i = 0; i < n; i++)
		a = a + d;
	return a;
}
