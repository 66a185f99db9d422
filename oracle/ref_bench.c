/* Timing driver of the reference renderer built as it ships (oracle/Makefile
 * target refbench: its own sources and flags, -O3, a static executable, no
 * position-independent code) - the CPU baseline of bench.py.  Test and
 * measurement infrastructure only.
 *
 *   ref_bench mesh    W H rings segs texture.png  BUDGET_S MIN MAX DUMP
 *   ref_bench tris    W H ntris seed              BUDGET_S MIN MAX DUMP
 *   ref_bench mix     W H ntris nlines npoints seed BUDGET_S MIN MAX DUMP
 *   ref_bench terrain W H gx gz pose              BUDGET_S MIN MAX DUMP
 *
 * Sets the workload up, draws one warm-up frame, then frames until BUDGET_S
 * seconds have passed (at least MIN, at most MAX frames), prints one JSON
 * line with every frame's time, and writes the last frame's colour, depth
 * and stencil arrays to DUMP. */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int srpBenchMeshSetup(int w, int h, int rings, int segs, const char* texPath);
int srpBenchTrisSetup(int w, int h, int nTri, uint64_t seed);
int srpBenchMixSetup(int w, int h, int nTri, int nLine, int nPoint, uint64_t seed);
int srpBenchTerrainSetup(int w, int h, int nx, int nz, int pose, int capture);
void srpBenchFrame(void);
void srpBenchRead(uint32_t* color, float* depth, uint8_t* stencil);
void srpBenchFinish(void);

static double now(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double) t.tv_sec + 1e-9 * (double) t.tv_nsec;
}

int main(int argc, char** argv)
{
	if (argc < 2)
		return 2;
	const char* kind = argv[1];
	int rc, w, h, a = 0;
	if (!strcmp(kind, "mesh") && argc == 11)
	{
		w = atoi(argv[2]); h = atoi(argv[3]);
		rc = srpBenchMeshSetup(w, h, atoi(argv[4]), atoi(argv[5]), argv[6]);
		a = 7;
	}
	else if (!strcmp(kind, "tris") && argc == 10)
	{
		w = atoi(argv[2]); h = atoi(argv[3]);
		rc = srpBenchTrisSetup(w, h, atoi(argv[4]), strtoull(argv[5], NULL, 0));
		a = 6;
	}
	else if (!strcmp(kind, "mix") && argc == 12)
	{
		w = atoi(argv[2]); h = atoi(argv[3]);
		rc = srpBenchMixSetup(w, h, atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), strtoull(argv[7], NULL, 0));
		a = 8;
	}
	else if (!strcmp(kind, "terrain") && argc == 11)
	{
		w = atoi(argv[2]); h = atoi(argv[3]);
		rc = srpBenchTerrainSetup(w, h, atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), 0);
		a = 7;
	}
	else
	{
		fprintf(stderr, "ref_bench: bad arguments\n");
		return 2;
	}
	if (rc != 0)
	{
		fprintf(stderr, "ref_bench: setup failed (%d)\n", rc);
		return 1;
	}
	const double budget = atof(argv[a]);
	const int minFrames = atoi(argv[a + 1]), maxFrames = atoi(argv[a + 2]);
	const char* dump = argv[a + 3];
	srpBenchFrame();   /* warm-up */
	double* t = (double*) malloc(sizeof(double) * (size_t) (maxFrames > 0 ? maxFrames : 1));
	int n = 0;
	const double t0 = now();
	while (n < maxFrames && (n < minFrames || now() - t0 < budget))
	{
		const double s = now();
		srpBenchFrame();
		t[n++] = now() - s;
	}
	printf("{\"frames\": %d, \"frame_s\": [", n);
	for (int i = 0; i < n; i++)
		printf("%s%.9f", i ? ", " : "", t[i]);
	printf("]}\n");
	const size_t px = (size_t) w * (size_t) h;
	uint32_t* c = (uint32_t*) malloc(px * 4);
	float* d = (float*) malloc(px * 4);
	uint8_t* st = (uint8_t*) malloc(px);
	srpBenchRead(c, d, st);
	FILE* f = fopen(dump, "wb");
	if (!f)
		return 1;
	fwrite(c, 4, px, f);
	fwrite(d, 4, px, f);
	fwrite(st, 1, px, f);
	fclose(f);
	srpBenchFinish();
	free(t); free(c); free(d); free(st);
	return 0;
}
