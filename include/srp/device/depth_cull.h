/* Conservative depth range of a triangle over one 8x8 tile, for tile-level
 * depth culling in the rasteriser.
 *
 * The reference computes a fragment's depth at pixel (x, y) of triangle T as
 *   d = fl(fl(fl(zw0*l0) + fl(zw1*l1)) + fl(zw2*l2))        (interpolation.c:44-46)
 * where each l_e is the float sum l0_e + dy_e + ... + dy_e + dx_e + ... + dx_e
 * (ny = y - y0 additions of dy, then nx = x - x0 of dx; triangle.c:101-109).
 * In exact arithmetic the same inputs give the affine plane
 *   D(nx, ny) = sum_e zw_e * (l0_e + ny*dy_e + nx*dx_e).
 * Recursive summation of n terms errs by at most gamma(n-1) * sum|terms|
 * (gamma(k) = k*u / (1 - k*u), u = 2^-24), so |l_e - L_e| <= E_e =
 * gamma(nx + ny) * M_e with M_e = |l0_e| + ny*|dy_e| + nx*|dx_e|; the
 * three-term dot product adds at most gamma(3) * sum|zw_e * l_e|.  Hence for
 * every pixel of the tile inside T's box
 *   |d - D| <= sum_e |zw_e| * (E_e + gamma(3) * (M_e + E_e)),
 * taking nx, ny at the far corner of the tile/box overlap (both bounds grow
 * with them).  D is affine, so its extremes over that rectangle lie at its
 * corners; they are evaluated in double, whose own rounding (< 1e-15
 * relative) is covered by a further 1e-12 * sum|zw_e| * M_e.
 *
 * Host and device code share this header (tests/native/depth_cull_check.c
 * checks the range against every pixel's sequential walk); FP contraction
 * must be off. */
#pragma once

#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include "srp/context.h"
#include "srp/device/abi.h"
#include "srp/detail/qualifiers.h"

/* [*lo, *hi] holds the depth of every fragment T can have in the 8x8 tile
 * (tx, ty).  False when the tile and T's box do not overlap or an input is
 * not finite (no range). */
SRP_HD static inline bool srpTileDepthRange(const SrpTri* T, uint32_t tx, uint32_t ty, double* lo, double* hi)
{
	const int bx0 = (int) (tx * 8u), by0 = (int) (ty * 8u);
	const int xs = bx0 > (int) T->x0 ? bx0 : (int) T->x0;
	const int ys = by0 > (int) T->y0 ? by0 : (int) T->y0;
	const int xe = (bx0 + 7 < (int) T->x1 - 1) ? bx0 + 7 : (int) T->x1 - 1;
	const int ye = (by0 + 7 < (int) T->y1 - 1) ? by0 + 7 : (int) T->y1 - 1;
	if (xs > xe || ys > ye)
		return false;
	const double nx0 = (double) (xs - (int) T->x0), nx1 = (double) (xe - (int) T->x0);
	const double ny0 = (double) (ys - (int) T->y0), ny1 = (double) (ye - (int) T->y0);
	const double u = 0x1p-24;
	const double n = nx1 + ny1;
	const double gN = n * u / (1.0 - n * u);
	const double g3 = 3.0 * u / (1.0 - 3.0 * u);
	double c = 0.0, gx = 0.0, gy = 0.0, err = 0.0;
	for (int e = 0; e < 3; e++)
	{
		const double a = (double) T->zw[e], l = (double) T->l0[e], dx = (double) T->dx[e], dy = (double) T->dy[e];
		c += a * l;
		gx += a * dx;
		gy += a * dy;
		const double m = fabs(l) + ny1 * fabs(dy) + nx1 * fabs(dx);
		const double em = gN * m;
		err += fabs(a) * (em + g3 * (m + em) + 1e-12 * m);
	}
	*lo = c + fmin(nx0 * gx, nx1 * gx) + fmin(ny0 * gy, ny1 * gy) - err;
	*hi = c + fmax(nx0 * gx, nx1 * gx) + fmax(ny0 * gy, ny1 * gy) + err;
	return isfinite(*lo) && isfinite(*hi);
}

/* True when every fragment T could have in tile (tx, ty) fails the depth
 * compare `op` against every stored depth of the tile, whose range is
 * [zMin, zMax]: d >= lo > zMax for LESS/LEQUAL, d <= hi < zMin for
 * GREATER/GEQUAL, and always for NEVER.  NaN depths never cull (the
 * comparisons with them are false). */
SRP_HD static inline bool srpTileDepthCull(const SrpTri* T, uint32_t tx, uint32_t ty, uint32_t op, float zMin,
                                          float zMax)
{
	if (op == SRP_COMPARE_NEVER)
		return true;
	double lo, hi;
	if (!srpTileDepthRange(T, tx, ty, &lo, &hi))
		return false;
	if (op == SRP_COMPARE_LESS || op == SRP_COMPARE_LEQUAL)
		return lo > (double) zMax;
	if (op == SRP_COMPARE_GREATER || op == SRP_COMPARE_GEQUAL)
		return hi < (double) zMin;
	return false;
}
