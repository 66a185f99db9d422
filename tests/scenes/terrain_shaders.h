/* Shaders of the terrain workload: a fBm heightfield lit by four point
 * lights.  The plain fragment shader is the expensive procedural one
 * (8-octave fBm of 3D gradient noise for the albedo, Blinn-Phong over four
 * point lights); the capture shader writes two 16-bit surface features per
 * pass into RGBA8 so that four passes through the ordinary pipeline yield the
 * G-buffer (position, normal, UV) of exactly the fragments that survive the
 * depth test.
 *
 * Everything is plain IEEE float arithmetic with correctly rounded sqrt and
 * division, so the host (oracle) and device builds agree bit for bit.
 * Included by shaders.h. */
#pragma once

typedef struct VtxTerrain { vec3 position; vec3 normal; vec2 uv; } VtxTerrain;
typedef struct VarTerrain { vec3 position; vec3 normal; vec2 uv; } VarTerrain;

/* Normalisation of the eight surface features to [0,1] (position) and
 * [-1,1] (normal); the same affine map feeds the network. */
typedef struct TerrainBounds { float lo[3], inv[3]; } TerrainBounds;

typedef struct UniTerrain
{
	mat4 viewProj;
	TerrainBounds bounds;
	uint32_t pass;          /* capture pass 0..3 */
	uint32_t pad;
	const void* network;    /* device weights of the neural shader */
} UniTerrain;

#define TERRAIN_OCTAVES 8

SCENE_SHADER void vsTerrain(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxTerrain* v = (const VtxTerrain*) in->vertex;
	const UniTerrain* u = (const UniTerrain*) in->uniform;
	scStoreClip(out, mat4MultiplyVec4(&u->viewProj, VEC4_FROM_VEC3(v->position, 1.)));
	VarTerrain* o = (VarTerrain*) out->varyings;
	o->position = v->position;
	o->normal = v->normal;
	o->uv = v->uv;
}

/* Integer lattice hash (xor-shift-multiply finaliser). */
SCENE_SHADER uint32_t tnHash(int32_t x, int32_t y, int32_t z)
{
	uint32_t h = ((uint32_t) x * 0x8DA6B343u) ^ ((uint32_t) y * 0xD8163841u) ^ ((uint32_t) z * 0xCB1AB31Fu);
	h ^= h >> 16;
	h *= 0x7FEB352Du;
	h ^= h >> 15;
	h *= 0x846CA68Bu;
	h ^= h >> 16;
	return h;
}

/* Gradient of improved Perlin noise: 12 cube-edge directions (+4 repeats). */
SCENE_SHADER float tnGrad(uint32_t h, float x, float y, float z)
{
	const uint32_t k = h & 15u;
	const float u = k < 8u ? x : y;
	const float v = k < 4u ? y : ((k == 12u || k == 14u) ? x : z);
	return ((k & 1u) ? -u : u) + ((k & 2u) ? -v : v);
}

SCENE_SHADER float tnFade(float t)
{
	return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

SCENE_SHADER float tnLerp(float a, float b, float t)
{
	return a + t * (b - a);
}

/* 3D gradient noise, roughly in [-1,1]. */
SCENE_SHADER float tnNoise(float x, float y, float z)
{
	const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
	const int32_t X = (int32_t) fx0, Y = (int32_t) fy0, Z = (int32_t) fz0;
	const float dx = x - fx0, dy = y - fy0, dz = z - fz0;
	const float u = tnFade(dx), v = tnFade(dy), w = tnFade(dz);
	const float n000 = tnGrad(tnHash(X, Y, Z), dx, dy, dz);
	const float n100 = tnGrad(tnHash(X + 1, Y, Z), dx - 1.0f, dy, dz);
	const float n010 = tnGrad(tnHash(X, Y + 1, Z), dx, dy - 1.0f, dz);
	const float n110 = tnGrad(tnHash(X + 1, Y + 1, Z), dx - 1.0f, dy - 1.0f, dz);
	const float n001 = tnGrad(tnHash(X, Y, Z + 1), dx, dy, dz - 1.0f);
	const float n101 = tnGrad(tnHash(X + 1, Y, Z + 1), dx - 1.0f, dy, dz - 1.0f);
	const float n011 = tnGrad(tnHash(X, Y + 1, Z + 1), dx, dy - 1.0f, dz - 1.0f);
	const float n111 = tnGrad(tnHash(X + 1, Y + 1, Z + 1), dx - 1.0f, dy - 1.0f, dz - 1.0f);
	const float x00 = tnLerp(n000, n100, u), x10 = tnLerp(n010, n110, u);
	const float x01 = tnLerp(n001, n101, u), x11 = tnLerp(n011, n111, u);
	return tnLerp(tnLerp(x00, x10, v), tnLerp(x01, x11, v), w);
}

/* fBm: octaves of gradient noise, lacunarity 2, gain 1/2; each octave is
 * shifted so lattice points of successive octaves do not line up. */
SCENE_SHADER float tnFbm(float x, float y, float z, int octaves)
{
	float sum = 0.0f, amp = 0.5f;
	for (int k = 0; k < octaves; k++)
	{
		sum += amp * tnNoise(x, y, z);
		x = x * 2.0f + 17.13f;
		y = y * 2.0f + 3.71f;
		z = z * 2.0f + 9.29f;
		amp *= 0.5f;
	}
	return sum;
}

SCENE_SHADER float tnClamp01(float x)
{
	return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
}

SCENE_SHADER vec3 tnNormalize(vec3 v)
{
	const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
	return l > 0.0f ? VEC3(v.x / l, v.y / l, v.z / l) : v;
}

/* The procedural material: fBm albedo ramp (grass - rock - snow) lit by
 * four point lights, Blinn-Phong with exponent 32.  The viewer is at
 * infinity along a fixed direction, so the colour is a function of the
 * surface point (position, normal) alone. */
SCENE_SHADER void terrainShade(vec3 p, vec3 nrm, float out[4])
{
	const float n = tnFbm(p.x * 0.6f, p.y * 0.6f, p.z * 0.6f, TERRAIN_OCTAVES);
	const float t = tnClamp01(0.5f + 1.4f * n);
	const float snow = tnClamp01((p.y - 0.55f) * 2.5f + 0.8f * n);
	float alb[3];
	const float grass[3] = {0.20f, 0.36f, 0.12f}, rock[3] = {0.50f, 0.42f, 0.33f}, ice[3] = {0.90f, 0.92f, 0.96f};
	for (int c = 0; c < 3; c++)
		alb[c] = tnLerp(tnLerp(grass[c], rock[c], t), ice[c], snow);

	const vec3 N = tnNormalize(nrm);
	const vec3 V = VEC3(0.26726124f, 0.80178373f, 0.53452248f);
	const float lp[4][3] = {{-6.0f, 3.0f, -2.0f}, {5.0f, 2.5f, 3.0f}, {0.0f, 4.0f, 0.0f}, {7.0f, 3.0f, -4.0f}};
	const float lc[4][3] = {{1.00f, 0.85f, 0.70f}, {0.55f, 0.65f, 1.00f}, {0.90f, 0.90f, 0.90f}, {0.80f, 0.60f, 0.50f}};
	float c[3] = {0.08f * alb[0], 0.09f * alb[1], 0.11f * alb[2]};
	for (int k = 0; k < 4; k++)
	{
		vec3 L = VEC3(lp[k][0] - p.x, lp[k][1] - p.y, lp[k][2] - p.z);
		const float d2 = L.x * L.x + L.y * L.y + L.z * L.z;
		const float dl = sqrtf(d2);
		L = VEC3(L.x / dl, L.y / dl, L.z / dl);
		const float att = 1.0f / (1.0f + 0.04f * d2);
		const float ndl = N.x * L.x + N.y * L.y + N.z * L.z;
		const float diff = ndl > 0.0f ? ndl : 0.0f;
		const vec3 H = tnNormalize(VEC3(L.x + V.x, L.y + V.y, L.z + V.z));
		const float ndh = N.x * H.x + N.y * H.y + N.z * H.z;
		float s = ndh > 0.0f ? ndh : 0.0f;
		s = s * s; s = s * s; s = s * s; s = s * s; s = s * s;   /* ^32 */
		const float spec = diff > 0.0f ? 0.25f * s : 0.0f;
		for (int ch = 0; ch < 3; ch++)
			c[ch] += lc[k][ch] * att * (alb[ch] * diff + spec);
	}
	for (int ch = 0; ch < 3; ch++)
		out[ch] = c[ch] * 2.2f;
	out[3] = 1.0f;
}

SCENE_SHADER void fsTerrain(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarTerrain* v = (const VarTerrain*) in->varyings;
	terrainShade(v->position, v->normal, out->color);
}

/* The eight network inputs of a fragment: position mapped to [0,1] by the
 * terrain bounds, normal as is, UV. */
SCENE_SHADER void terrainFeatures(const VarTerrain* v, const TerrainBounds* b, float f[8])
{
	f[0] = (v->position.x - b->lo[0]) * b->inv[0];
	f[1] = (v->position.y - b->lo[1]) * b->inv[1];
	f[2] = (v->position.z - b->lo[2]) * b->inv[2];
	f[3] = v->normal.x;
	f[4] = v->normal.y;
	f[5] = v->normal.z;
	f[6] = v->uv.x;
	f[7] = v->uv.y;
}

/* Capture pass k writes features 2k and 2k+1 as 16-bit fixed point
 * (position and UV over [0,1], normal over [-1,1]): R,G = high,low byte of
 * the first, B,A of the second.  A byte b is written as (b+0.5)/255, which
 * the colour packing (x*255, clamp, truncate) turns back into b. */
SCENE_SHADER void fsTerrainCapture(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarTerrain* v = (const VarTerrain*) in->varyings;
	const UniTerrain* u = (const UniTerrain*) in->uniform;
	float f[8];
	terrainFeatures(v, &u->bounds, f);
	for (int j = 0; j < 2; j++)
	{
		const uint32_t k = 2u * u->pass + (uint32_t) j;
		float x = f[k];
		if (k >= 3u && k <= 5u)
			x = 0.5f * x + 0.5f;
		const float q = tnClamp01(x) * 65535.0f + 0.5f;
		const uint32_t qi = (uint32_t) q;
		out->color[2 * j + 0] = ((float) (qi >> 8) + 0.5f) / 255.0f;
		out->color[2 * j + 1] = ((float) (qi & 255u) + 0.5f) / 255.0f;
	}
}
