/* Host check of srpTileDepthRange (include/srp/device/depth_cull.h): for
 * random and adversarial triangles and tiles, every pixel of the tile inside
 * the triangle's box gets the depth the reference computes (barycentrics by
 * the sequential walk of raster/triangle.c:101-109, depth by the three-term
 * sum of interpolation.c:44-46), and that depth must lie in [lo, hi].
 * Prints the number of violations; exit code 1 if any. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "srp/device/depth_cull.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void)
{
	uint64_t z = (st += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static double u01(void) { return (double) (rnd() >> 11) * (1.0 / 9007199254740992.0); }
static int irange(int a, int b) { return a + (int) (rnd() % (uint64_t) (b - a + 1)); }
static float logU(double lo, double hi) { return (float) pow(10.0, lo + (hi - lo) * u01()) * ((rnd() & 1) ? 1.0f : -1.0f); }

static long bad = 0, cases = 0, pixels = 0;

static void check(const SrpTri* T, uint32_t tx, uint32_t ty)
{
	double lo, hi;
	if (!srpTileDepthRange(T, tx, ty, &lo, &hi))
		return;
	cases++;
	const int xs = (int) (tx * 8) > T->x0 ? (int) (tx * 8) : T->x0;
	const int ys = (int) (ty * 8) > T->y0 ? (int) (ty * 8) : T->y0;
	const int xe = ((int) (tx * 8 + 7) < T->x1 - 1) ? (int) (tx * 8 + 7) : T->x1 - 1;
	const int ye = ((int) (ty * 8 + 7) < T->y1 - 1) ? (int) (ty * 8 + 7) : T->y1 - 1;
	float row[3] = {T->l0[0], T->l0[1], T->l0[2]};
	for (int y = T->y0; y <= ye; y++)
	{
		if (y >= ys)
		{
			float l[3] = {row[0], row[1], row[2]};
			for (int x = T->x0; x <= xe; x++)
			{
				if (x >= xs)
				{
					const float d = T->zw[0] * l[0] + T->zw[1] * l[1] + T->zw[2] * l[2];
					pixels++;
					if (isfinite(d) && !((double) d >= lo && (double) d <= hi))
					{
						if (bad < 10)
							printf("violation: d=%.9g range [%.9g, %.9g] at (%d,%d) box (%d,%d)-(%d,%d)\n", d, lo, hi, x,
							       y, T->x0, T->y0, T->x1, T->y1);
						bad++;
					}
				}
				for (int e = 0; e < 3; e++) l[e] = l[e] + T->dx[e];
			}
		}
		for (int e = 0; e < 3; e++) row[e] = row[e] + T->dy[e];
	}
}

/* A triangle set up from screen-space vertices the way triangle.c:113-160
 * does (float edge function, barycentrics at the box's first pixel centre). */
static int realistic(SrpTri* T, int W, int H)
{
	float sx[3], sy[3];
	const double cx = u01() * W, cy = u01() * H, r = pow(10.0, 0.5 + 3.0 * u01());
	for (int i = 0; i < 3; i++)
	{
		sx[i] = (float) (cx + r * (2.0 * u01() - 1.0));
		sy[i] = (float) (cy + r * (2.0 * u01() - 1.0));
	}
	float ex[3], ey[3];
	for (int i = 0; i < 3; i++)
	{
		ex[i] = sx[(i + 1) % 3] - sx[i];
		ey[i] = sy[(i + 1) % 3] - sy[i];
	}
	const float area = fabsf(ex[0] * ey[2] - ey[0] * ex[2]);
	if (area <= 1e-3f)
		return 0;
	double mnx = fmin(sx[0], fmin(sx[1], sx[2])), mny = fmin(sy[0], fmin(sy[1], sy[2]));
	double mxx = fmax(sx[0], fmax(sx[1], sx[2])), mxy = fmax(sy[0], fmax(sy[1], sy[2]));
	const int x0 = (int) fmax(floor(mnx), 0.0), y0 = (int) fmax(floor(mny), 0.0);
	const int x1 = (int) fmin(ceil(mxx), (double) W), y1 = (int) fmin(ceil(mxy), (double) H);
	if (x1 <= x0 || y1 <= y0)
		return 0;
	const float px = (float) x0 + 0.5f, py = (float) y0 + 0.5f;
	for (int i = 0; i < 3; i++)
	{
		const int j = (i + 1) % 3;
		const float rx = px - sx[j], ry = py - sy[j];
		T->l0[i] = (rx * ey[j] - ry * ex[j]) / area;
		T->dx[i] = ey[j] / area;
		T->dy[i] = -ex[j] / area;
		T->zw[i] = (float) (2.0 * u01() - 1.0) * ((rnd() % 4 == 0) ? 50.0f : 1.0f);
	}
	T->x0 = (uint16_t) x0; T->y0 = (uint16_t) y0; T->x1 = (uint16_t) x1; T->y1 = (uint16_t) y1;
	return 1;
}

int main(int argc, char** argv)
{
	const long n = argc > 1 ? atol(argv[1]) : 20000;
	for (long k = 0; k < n; k++)
	{
		SrpTri T;
		memset(&T, 0, sizeof(T));
		const int kind = (int) (k % 3);
		if (kind == 0)
		{
			if (!realistic(&T, 2048, 2048))
				continue;
		}
		else
		{
			/* arbitrary finite inputs, walks up to a few thousand steps */
			T.x0 = (uint16_t) irange(0, 1500);
			T.y0 = (uint16_t) irange(0, 1500);
			T.x1 = (uint16_t) (T.x0 + irange(1, kind == 1 ? 64 : 2500));
			T.y1 = (uint16_t) (T.y0 + irange(1, kind == 1 ? 64 : 2500));
			for (int e = 0; e < 3; e++)
			{
				T.l0[e] = logU(-4, 2);
				T.dx[e] = (rnd() % 8 == 0) ? 0.0f : logU(-7, -0.5);
				T.dy[e] = (rnd() % 8 == 0) ? 0.0f : logU(-7, -0.5);
				T.zw[e] = logU(-3, 1.5);
			}
			if (rnd() % 4 == 0)
			{
				/* near-cancelling depth: zw_2 chosen so the plane is ~0 at the box origin */
				const float s = T.l0[2] != 0.0f ? T.l0[2] : 1.0f;
				T.zw[2] = -(T.zw[0] * T.l0[0] + T.zw[1] * T.l0[1]) / s;
			}
		}
		/* a tile near the box's far corner (long walks) and one anywhere in it */
		const uint32_t txa = (uint32_t) (T.x0 / 8), txb = (uint32_t) ((T.x1 - 1) / 8);
		const uint32_t tya = (uint32_t) (T.y0 / 8), tyb = (uint32_t) ((T.y1 - 1) / 8);
		check(&T, txb, tyb);
		check(&T, (uint32_t) irange((int) txa, (int) txb), (uint32_t) irange((int) tya, (int) tyb));
	}
	printf("%ld violations in %ld tiles (%ld pixels)\n", bad, cases, pixels);
	return bad ? 1 : 0;
}
