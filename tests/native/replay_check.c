/* Host check of srpReplay (include/srp/device/replay.h): equal, bit for
 * bit, to n sequential float additions, on random and adversarial cases
 * (steps of exactly half a spacing, zero crossings, binade boundaries,
 * tiny and huge steps).  Prints the number of mismatches; exit code 1 if
 * any. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "srp/device/replay.h"

static float naive(float a, float d, int n)
{
	for (int k = 0; k < n; k++)
		a = a + d;
	return a;
}

static uint64_t st = 0x243F6A8885A308D3ull;
static uint64_t rnd(void)
{
	uint64_t z = (st += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static double u01(void) { return (double) (rnd() >> 11) * (1.0 / 9007199254740992.0); }

static long bad = 0, total = 0;
static void check(float a, float d, int n)
{
	const float x = naive(a, d, n), y = srpReplay(a, d, n);
	total++;
	uint32_t bx, by;
	memcpy(&bx, &x, 4);
	memcpy(&by, &y, 4);
	if (bx != by && !(isnan(x) && isnan(y)))
	{
		if (bad < 10)
			printf("mismatch a=%a d=%a n=%d naive=%a fast=%a\n", a, d, n, x, y);
		bad++;
	}
}

int main(int argc, char** argv)
{
	const long iters = argc > 1 ? atol(argv[1]) : 200000;
	for (long i = 0; i < iters; i++)
	{
		const int kind = (int) (rnd() % 6);
		float a, d;
		int n = (int) (rnd() % 3000);
		if (kind == 0)
		{   /* barycentric-like: values around [-1, 2], small steps */
			a = (float) (u01() * 3.0 - 1.0);
			d = (float) ((u01() - 0.5) * pow(10.0, -1.0 - 5.0 * u01()));
		}
		else if (kind == 1)
		{   /* steps of exactly (m + 1/2) spacings of a's binade */
			a = (float) (ldexp(1.0 + u01(), (int) (rnd() % 20) - 10) * ((rnd() & 1) ? 1 : -1));
			int e;
			frexp(a, &e);
			const double u = ldexp(1.0, e - 24);
			const double m = (double) (rnd() % 64);
			d = (float) ((m + 0.5) * u * ((rnd() & 1) ? 1 : -1));
		}
		else if (kind == 2)
		{   /* sums that cross zero */
			a = (float) (u01() * 0.01);
			d = -(float) (u01() * 1e-4);
			n = (int) (rnd() % 400);
		}
		else if (kind == 3)
		{   /* start near a binade boundary */
			const int e = (int) (rnd() % 30) - 15;
			a = (float) ldexp(1.0, e);
			a = nextafterf(a, (rnd() & 1) ? 0.0f : 1e30f);
			d = (float) (ldexp(1.0, e - 20) * (u01() - 0.5) * 4.0);
		}
		else if (kind == 4)
		{   /* tiny steps that stop changing the sum, and huge ones */
			a = (float) (u01() * 100.0 - 50.0);
			d = (float) (ldexp(u01(), -(int) (rnd() % 40)) * ((rnd() & 1) ? 1 : -1));
			if (rnd() % 4 == 0) d = (float) (d * 1e6);
		}
		else
		{   /* anything */
			uint32_t ba = (uint32_t) rnd(), bd = (uint32_t) rnd();
			memcpy(&a, &ba, 4);
			memcpy(&d, &bd, 4);
			if (!isfinite(a) || !isfinite(d)) { a = 1.0f; d = 0.25f; }
			n = (int) (rnd() % 200);
		}
		check(a, d, n);
	}
	/* walks that land exactly on a power of two while the magnitude
	 * shrinks, or on the last value below one while it grows */
	for (long i = 0; i < iters / 4; i++)
	{
		const int e = (int) (rnd() % 40) - 20;
		const double u = ldexp(1.0, e - 23);
		const float sgn = (rnd() & 1) ? 1.0f : -1.0f;
		const int up = (int) (rnd() & 1);
		const double m = (double) (1 + rnd() % 16);
		float a, d;
		if (up)
		{
			a = sgn * (float) (ldexp(1.0, e + 1) - u * (double) (1 + rnd() % 64));
			d = sgn * (float) (u * m);
		}
		else
		{
			a = sgn * (float) (ldexp(1.0, e) + u * (double) (rnd() % 64));
			d = -sgn * (float) (u * m * ((rnd() % 3 == 0) ? 1.5 : 1.0));
		}
		check(a, d, (int) (30 + rnd() % 2000));
	}
	/* special values */
	check(0.0f, 0.1f, 100);
	check(-0.0f, 0.0f, 5);
	check(1.0f, 0.0f, 7);
	check(INFINITY, 1.0f, 3);
	check(1.0f, NAN, 3);
	check(1e-40f, 1e-41f, 1000);
	printf("%ld of %ld cases differ\n", bad, total);
	return bad != 0;
}
