/* Wavefront OBJ mesh loader for the test scenes and examples.
 *
 * Same behaviour as the reference's examples/utility/objparser.c:8-82:
 * `v`, `vt`, `vn` records parsed with "%f" (correctly rounded), faces of
 * exactly nine integers "f v/t/n v/t/n v/t/n" de-indexed into one new
 * vertex per face corner, indices 0,1,2,... in file order. */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <srp/vec.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SRPObjVertex {
	vec3 position;
	vec2 uv;
	vec3 normal;
} SRPObjVertex;

typedef struct SRPObjMesh {
	SRPObjVertex* vertices;
	size_t vertexCount;
	uint32_t* indices;
	size_t indexCount;
} SRPObjMesh;

bool srpObjLoad(const char* path, SRPObjMesh* mesh);
void srpObjFree(SRPObjMesh* mesh);

#ifdef __cplusplus
}
#endif
