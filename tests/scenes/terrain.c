/* The terrain workload: a seeded fBm heightfield of nx x nz cells
 * ((nx+1)(nz+1) vertices {position, normal, uv}, 2*nx*nz u32-indexed
 * triangles) over x in [-10,10], z in [-5,5], seen from seeded camera poses
 * (fov_y 60 deg, near 0.1, far 100), depth LESS, back faces culled.
 *
 * Pose k uses seed 0x5EED0005 + k.  Poses 0..255 are the training poses of
 * the neural shader and 256..319 the held-out evaluation poses.
 *
 * Written against the public API only (links against the oracle, the
 * compiled reference and the GPU library). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "shaders.h"
#include "scenes.h"
#include "terrain.h"

#ifndef M_PI
#  define M_PI 3.14159265358979323846
#endif

#define TERRAIN_X0 (-10.0)
#define TERRAIN_X1 (10.0)
#define TERRAIN_Z0 (-5.0)
#define TERRAIN_Z1 (5.0)

static uint64_t trNext(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static double trU(uint64_t* s) { return (double) (trNext(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Terrain height at world (x, z): 6 octaves of the shaders' gradient noise. */
static double trHeight(double x, double z)
{
	return 1.6 * (double) tnFbm((float) (x * 0.16), 0.37f, (float) (z * 0.16), 6) + 0.25;
}

void srpTerrainBuild(int nx, int nz, VtxTerrain** verts, size_t* nVerts, uint32_t** idx, size_t* nIdx,
                     TerrainBounds* bounds)
{
	const size_t nv = (size_t) (nx + 1) * (size_t) (nz + 1);
	VtxTerrain* v = (VtxTerrain*) malloc(nv * sizeof(VtxTerrain));
	const double dx = (TERRAIN_X1 - TERRAIN_X0) / nx, dz = (TERRAIN_Z1 - TERRAIN_Z0) / nz;
	double ylo = 1e30, yhi = -1e30;
	for (int j = 0; j <= nz; j++)
		for (int i = 0; i <= nx; i++)
		{
			const double x = TERRAIN_X0 + i * dx, z = TERRAIN_Z0 + j * dz;
			const double h = trHeight(x, z);
			/* central differences of the height function at a fixed step */
			const double e = 0.01;
			const double hx = (trHeight(x + e, z) - trHeight(x - e, z)) / (2 * e);
			const double hz = (trHeight(x, z + e) - trHeight(x, z - e)) / (2 * e);
			const double l = sqrt(hx * hx + 1.0 + hz * hz);
			VtxTerrain* p = &v[(size_t) j * (nx + 1) + i];
			p->position = VEC3((float) x, (float) h, (float) z);
			p->normal = VEC3((float) (-hx / l), (float) (1.0 / l), (float) (-hz / l));
			p->uv = VEC2((float) i / (float) nx, (float) j / (float) nz);
			if (h < ylo) ylo = h;
			if (h > yhi) yhi = h;
		}
	const size_t ni = (size_t) nx * (size_t) nz * 6;
	uint32_t* ix = (uint32_t*) malloc(ni * sizeof(uint32_t));
	size_t k = 0;
	for (int j = 0; j < nz; j++)
		for (int i = 0; i < nx; i++)
		{
			const uint32_t a = (uint32_t) (j * (nx + 1) + i), b = a + 1;
			const uint32_t c = a + (uint32_t) (nx + 1), d = c + 1;
			ix[k++] = a; ix[k++] = b; ix[k++] = c;
			ix[k++] = b; ix[k++] = d; ix[k++] = c;
		}
	bounds->lo[0] = (float) TERRAIN_X0;
	bounds->lo[1] = (float) ylo;
	bounds->lo[2] = (float) TERRAIN_Z0;
	bounds->inv[0] = (float) (1.0 / (TERRAIN_X1 - TERRAIN_X0));
	bounds->inv[1] = (float) (1.0 / (yhi - ylo));
	bounds->inv[2] = (float) (1.0 / (TERRAIN_Z1 - TERRAIN_Z0));
	*verts = v; *nVerts = nv; *idx = ix; *nIdx = ni;
}

/* Camera of pose k: eye above the terrain, looking down 20-45 degrees. */
void srpTerrainPose(int pose, int width, int height, UniTerrain* u)
{
	uint64_t s = 0x5EED0005ull + (uint64_t) pose;
	const double ex = -7.0 + 14.0 * trU(&s), ez = -3.5 + 7.0 * trU(&s);
	const double ey = trHeight(ex, ez) + 1.2 + 2.3 * trU(&s);
	const double yaw = 2.0 * M_PI * trU(&s), pitch = (20.0 + 25.0 * trU(&s)) * M_PI / 180.0;
	const double f[3] = {cos(pitch) * cos(yaw), -sin(pitch), cos(pitch) * sin(yaw)};
	/* right = up x forward, up' = forward x right (view space: x right, y up, z forward) */
	double r[3] = {f[2], 0.0, -f[0]};
	const double rl = sqrt(r[0] * r[0] + r[2] * r[2]);
	r[0] /= rl; r[2] /= rl;
	const double up[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
	const double e[3] = {ex, ey, ez};
	mat4 view;
	memset(&view, 0, sizeof(view));
	const double* rows[3] = {r, up, f};
	for (int i = 0; i < 3; i++)
	{
		for (int j = 0; j < 3; j++)
			view.data[i][j] = (float) rows[i][j];
		view.data[i][3] = (float) -(rows[i][0] * e[0] + rows[i][1] * e[1] + rows[i][2] * e[2]);
	}
	view.data[3][3] = 1.0f;
	const float t = (float) tan(M_PI / 6.0), n = 0.1f, fa = 100.f;
	const float aspect = (float) width / (float) height;
	const mat4 proj = mat4ConstructPerspectiveProjection(-t * n * aspect, t * n * aspect, -t * n, t * n, n, fa);
	u->viewProj = mat4MultiplyMat4(&proj, &view);
	u->pass = 0;
	u->pad = 0;
	u->network = NULL;
}

void srpTerrainState(void)
{
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LESS);
	srpRasterCullFace(SRP_FACE_BACK);
}

SRPVaryingInfo g_terrainVaryings[3] = {
	{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE},
	{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE},
	{2, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE},
};
