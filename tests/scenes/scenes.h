/* Scene runner interface shared by the CPU and GPU scene modules. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SceneArgs
{
	const char* resDir;   /* directory holding utah_teapot.obj, stoneWall.png */
	const char* texPath;  /* texture file of cfg/mesh */
	int width, height;    /* framebuffer of the cfg scenes (0 = default) */
	int p0, p1, p2;       /* scene-specific sizes */
	float f0;             /* scene-specific parameter */
	uint64_t seed;
} SceneArgs;

typedef struct SceneOut
{
	uint32_t* color;   /* caller buffers of `capacity` pixels */
	float* depth;
	uint8_t* stencil;
	size_t capacity;
	size_t width, height;
} SceneOut;

int srpSceneCount(void);
const char* srpSceneName(int i);
/** 0 on success, -1 unknown scene, -2 output too small, -3 missing resource. */
int srpSceneRun(const char* name, const SceneArgs* args, SceneOut* out);
int srpSceneFuzz(const SceneArgs* args, SceneOut* out);

#ifdef __cplusplus
}
#endif
