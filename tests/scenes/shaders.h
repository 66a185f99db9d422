/* Shaders of the test scenes and benchmark workloads.
 *
 * Written once in the common subset of C and HIP C++: the CPU builds (the
 * oracle and the compiled reference) wrap them as host function pointers
 * (host_shaders.c); the GPU build registers them as device shaders
 * (device_shaders.hip).  The uniform and vertex layouts below are shared by
 * the scene setup code (scenes.c).
 *
 * The reference scene programs (tests/scenes/<group>/<name>.c of kitrofimov/srp) define
 * the inputs that these shaders restate: same vertex formats, same matrix
 * products in the same order, same colour formulas. */
#pragma once

#define SRP_INCLUDE_VEC
#define SRP_INCLUDE_MAT
#include <math.h>
#include <srp/srp.h>

#if defined(__HIPCC__)
#  define SCENE_SHADER static __host__ __device__ inline
#else
#  define SCENE_SHADER static inline
#endif

/* ------------------------------------------------------------ layouts */

typedef struct VtxPosColor { vec3 position; vec3 color; } VtxPosColor;
typedef struct VtxPosUV { vec3 position; vec2 uv; } VtxPosUV;
typedef struct VtxPosByte { vec3 position; uint8_t color; } VtxPosByte;
typedef struct VtxObj { vec3 position; vec2 uv; vec3 normal; } VtxObj;

typedef struct UniMVP { mat4 model; mat4 view; mat4 projection; } UniMVP;
typedef struct UniFrameMVP { size_t frameCount; mat4 model; mat4 view; mat4 projection; } UniFrameMVP;
typedef struct UniFrameMVPTex { size_t frameCount; mat4 model; mat4 view; mat4 projection; SRPTexture* texture; } UniFrameMVPTex;
typedef struct UniMat { mat4 m; } UniMat;

typedef struct VarColor { vec3 color; } VarColor;
typedef struct VarUV { vec2 uv; } VarUV;
typedef struct VarByte { uint8_t color; } VarByte;

SCENE_SHADER vec4 scMVP(const mat4* m, const mat4* v, const mat4* p, vec3 pos)
{
	vec4 r = VEC4_FROM_VEC3(pos, 1.);
	r = mat4MultiplyVec4(m, r);
	r = mat4MultiplyVec4(v, r);
	r = mat4MultiplyVec4(p, r);
	return r;
}

SCENE_SHADER void scStoreClip(SRPVertexShaderOut* out, vec4 p)
{
	out->clipPosition[0] = p.x; out->clipPosition[1] = p.y;
	out->clipPosition[2] = p.z; out->clipPosition[3] = p.w;
}

/* ------------------------------------------------------- vertex shaders */

/* position,colour -> model/view/projection; colour varying */
SCENE_SHADER void vsColorMVP(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosColor* v = (const VtxPosColor*) in->vertex;
	const UniMVP* u = (const UniMVP*) in->uniform;
	scStoreClip(out, scMVP(&u->model, &u->view, &u->projection, v->position));
	((VarColor*) out->varyings)->color = v->color;
}

/* position -> model/view/projection, no varyings */
SCENE_SHADER void vsPosMVP(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosColor* v = (const VtxPosColor*) in->vertex;
	const UniMVP* u = (const UniMVP*) in->uniform;
	scStoreClip(out, scMVP(&u->model, &u->view, &u->projection, v->position));
}

/* position times one matrix; colour varying */
SCENE_SHADER void vsColorMat(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosColor* v = (const VtxPosColor*) in->vertex;
	const UniMat* u = (const UniMat*) in->uniform;
	scStoreClip(out, mat4MultiplyVec4(&u->m, VEC4_FROM_VEC3(v->position, 1.)));
	((VarColor*) out->varyings)->color = v->color;
}

/* position passed through (w = 1); colour varying */
SCENE_SHADER void vsColorPass(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosColor* v = (const VtxPosColor*) in->vertex;
	scStoreClip(out, VEC4_FROM_VEC3(v->position, 1.));
	((VarColor*) out->varyings)->color = v->color;
}

/* position passed through, no varyings */
SCENE_SHADER void vsPosPass(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosColor* v = (const VtxPosColor*) in->vertex;
	scStoreClip(out, VEC4_FROM_VEC3(v->position, 1.));
}

/* OBJ vertex -> model/view/projection, no varyings */
SCENE_SHADER void vsObjMVP(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxObj* v = (const VtxObj*) in->vertex;
	const UniFrameMVP* u = (const UniFrameMVP*) in->uniform;
	scStoreClip(out, scMVP(&u->model, &u->view, &u->projection, v->position));
}

/* position,uv -> model/view/projection; uv varying */
SCENE_SHADER void vsUVMVP(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosUV* v = (const VtxPosUV*) in->vertex;
	const UniFrameMVPTex* u = (const UniFrameMVPTex*) in->uniform;
	scStoreClip(out, scMVP(&u->model, &u->view, &u->projection, v->position));
	((VarUV*) out->varyings)->uv = v->uv;
}

/* position,byte -> one matrix; byte varying */
SCENE_SHADER void vsByteMat(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosByte* v = (const VtxPosByte*) in->vertex;
	const UniMat* u = (const UniMat*) in->uniform;
	scStoreClip(out, mat4MultiplyVec4(&u->m, VEC4_FROM_VEC3(v->position, 1.)));
	((VarByte*) out->varyings)->color = v->color;
}

/* ----------------------------------------------------- fragment shaders */

SCENE_SHADER void fsColor(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarColor* i = (const VarColor*) in->varyings;
	out->color[0] = i->color.x;
	out->color[1] = i->color.y;
	out->color[2] = i->color.z;
	out->color[3] = 1.;
}

SCENE_SHADER void fsWhite(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	(void) in;
	out->color[0] = 1.; out->color[1] = 1.; out->color[2] = 1.; out->color[3] = 1.;
}

SCENE_SHADER void fsRed(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	(void) in;
	out->color[0] = 1; out->color[1] = 0; out->color[2] = 0; out->color[3] = 1;
}

/* colour from the primitive ID (reference scenes clipping/triangle.c et al.) */
SCENE_SHADER void fsPrimID(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const int id = (int) in->primitiveID;
	out->color[0] = (float) (((id * 97) % 255) / 255.);
	out->color[1] = (float) (((id * 57) % 255) / 255.);
	out->color[2] = (float) (((id * 23) % 255) / 255.);
	out->color[3] = 1.;
}

SCENE_SHADER void fsPrimIDHalf(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const int id = (int) in->primitiveID;
	out->color[0] = (float) (((id * 97) % 255) / 255. / 2.);
	out->color[1] = (float) (((id * 57) % 255) / 255. / 2.);
	out->color[2] = (float) (((id * 23) % 255) / 255. / 2.);
	out->color[3] = 1.;
}

SCENE_SHADER void fsTexture(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarUV* i = (const VarUV*) in->varyings;
	const UniFrameMVPTex* u = (const UniFrameMVPTex*) in->uniform;
	srpTextureGetFilteredColor(u->texture, i->uv.x, i->uv.y, out->color);
}

SCENE_SHADER void fsByte(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarByte* v = (const VarByte*) in->varyings;
	float c[4] = {0, 0, 0, 0};
	if (v->color == 0) { c[0] = 1; c[3] = 1; }
	else if (v->color == 1) { c[1] = 1; c[3] = 1; }
	else if (v->color == 2) { c[2] = 1; c[3] = 1; }
	else return;
	for (int k = 0; k < 4; k++) out->color[k] = c[k];
}

/* ----------------------------------------- benchmark workload shaders */

/* Indexed textured mesh: perspective UV + flat u32 face tag. */
typedef struct VarUVTag { vec2 uv; uint32_t tag; } VarUVTag;

SCENE_SHADER void vsMeshTex(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxPosUV* v = (const VtxPosUV*) in->vertex;
	const UniFrameMVPTex* u = (const UniFrameMVPTex*) in->uniform;
	scStoreClip(out, scMVP(&u->model, &u->view, &u->projection, v->position));
	VarUVTag* o = (VarUVTag*) out->varyings;
	o->uv = v->uv;
	o->tag = (uint32_t) in->vertexID;
}

SCENE_SHADER void fsMeshTex(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarUVTag* i = (const VarUVTag*) in->varyings;
	const UniFrameMVPTex* u = (const UniFrameMVPTex*) in->uniform;
	float t[4];
	srpTextureGetFilteredColor(u->texture, i->uv.x, i->uv.y, t);
	const float shade = 1.0f - (float) (i->tag % 7u) * 0.03f;
	out->color[0] = t[0] * shade;
	out->color[1] = t[1] * shade;
	out->color[2] = t[2] * shade;
	out->color[3] = 1.0f;
}

/* Primitive mix: clip-space positions given directly; affine vec2 +
 * flat float + flat u32 varyings. */
typedef struct VtxMix { vec4 clip; vec2 st; float level; uint32_t tag; } VtxMix;
typedef struct VarMix { vec2 st; float level; uint32_t tag; } VarMix;

SCENE_SHADER void vsMix(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxMix* v = (const VtxMix*) in->vertex;
	scStoreClip(out, v->clip);
	VarMix* o = (VarMix*) out->varyings;
	o->st = v->st;
	o->level = v->level;
	o->tag = v->tag;
}

SCENE_SHADER void fsMix(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarMix* i = (const VarMix*) in->varyings;
	out->color[0] = i->st.x;
	out->color[1] = i->st.y;
	out->color[2] = i->level;
	out->color[3] = (float) (i->tag & 255u) / 255.0f;
}

/* Randomised scenes: every varying type, random interpolation modes. */
#pragma pack(push, 1)
typedef struct VarFuzz { float f[3]; double d; int32_t i32; uint16_t u16; int8_t i8; uint64_t u64; float g; } VarFuzz;
typedef struct VtxFuzz { vec4 clip; VarFuzz var; } VtxFuzz;
#pragma pack(pop)

SCENE_SHADER void vsFuzz(SRPVertexShaderIn* in, SRPVertexShaderOut* out)
{
	const VtxFuzz* v = (const VtxFuzz*) in->vertex;
	scStoreClip(out, v->clip);
	*(VarFuzz*) out->varyings = v->var;
}

SCENE_SHADER void scFuzzColor(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarFuzz* i = (const VarFuzz*) in->varyings;
	const uint32_t bits = (uint32_t) i->i32 ^ (uint32_t) i->u16 ^ (uint32_t) (int32_t) i->i8 ^ (uint32_t) i->u64;
	out->color[0] = i->f[0] * 0.5f + (float) (in->primitiveID % 5u) * 0.1f;
	out->color[1] = (float) (i->d * 0.25) + i->f[1] * i->g;
	out->color[2] = (float) (bits & 255u) / 255.0f;
	out->color[3] = in->frontFacing ? in->fragCoord[2] : in->fragCoord[3] * 0.01f;
}

SCENE_SHADER void fsFuzz(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	scFuzzColor(in, out);
}

/* Late depth: writes fragDepth, NaN (= keep the interpolated depth) for some fragments. */
SCENE_SHADER void fsFuzzDepth(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const VarFuzz* i = (const VarFuzz*) in->varyings;
	scFuzzColor(in, out);
	out->fragDepth = (i->f[2] > 0.5f) ? NAN : i->f[2] * 2.0f - 1.0f;
}

#include "terrain_shaders.h"
