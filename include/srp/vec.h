/* vec2/vec3/vec4.  Interface of reference include/srp/vec.h:15-104: packed
 * unions with the same members and swizzle views, same function set. */
#pragma once

#include <stdint.h>
#include "srp/detail/qualifiers.h"

#ifdef __cplusplus
#  pragma clang diagnostic push
#  pragma clang diagnostic ignored "-Wgnu-anonymous-struct"
#  pragma clang diagnostic ignored "-Wnested-anon-types"
#endif

#pragma pack(push, 1)

typedef union vec2 {
	struct { float x, y; };
	float v[2];
} vec2;

typedef union vec3 {
	struct { float x, y, z; };
	struct { vec2 xy; float _z; };
	struct { float _x; vec2 yz; };
	float v[3];
} vec3;

typedef union vec4 {
	struct { float x, y, z, w; };
	struct { vec2 xy; float _z, _w; };
	struct { float _x; vec2 yz; float __w; };
	struct { float __x, _y; vec2 zw; };
	struct { vec3 xyz; float ___w; };
	struct { float ___x; vec3 yzw; };
	float v[4];
} vec4;

#pragma pack(pop)

#ifdef __cplusplus
#  pragma clang diagnostic pop
#endif

#ifndef __cplusplus
#  define VEC2(x, y)           ((vec2) {{x, y}})
#  define VEC3(x, y, z)        ((vec3) {{x, y, z}})
#  define VEC4(x, y, z, w)     ((vec4) {{x, y, z, w}})
#  define VEC4_FROM_VEC3(v, a) ((vec4) {.xyz = (v), .___w = (a)})
#  define SWZ2(v, a, b)        ((vec2) {{(v).a, (v).b}})
#  define SWZ3(v, a, b, c)     ((vec3) {{(v).a, (v).b, (v).c}})
#  define SWZ4(v, a, b, c, d)  ((vec4) {{(v).a, (v).b, (v).c, (v).d}})
#else
/* C++ (HIP shader modules): same spelling, built by helper constructors. */
SRP_HD static inline vec2 srpMakeVec2(float x, float y) { vec2 r; r.x = x; r.y = y; return r; }
SRP_HD static inline vec3 srpMakeVec3(float x, float y, float z) { vec3 r; r.x = x; r.y = y; r.z = z; return r; }
SRP_HD static inline vec4 srpMakeVec4(float x, float y, float z, float w) { vec4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
SRP_HD static inline vec4 srpMakeVec4From3(vec3 v, float w) { vec4 r; r.x = v.x; r.y = v.y; r.z = v.z; r.w = w; return r; }
#  define VEC2(x, y)           srpMakeVec2((float) (x), (float) (y))
#  define VEC3(x, y, z)        srpMakeVec3((float) (x), (float) (y), (float) (z))
#  define VEC4(x, y, z, w)     srpMakeVec4((float) (x), (float) (y), (float) (z), (float) (w))
#  define VEC4_FROM_VEC3(v, a) srpMakeVec4From3((v), (float) (a))
#  define SWZ2(v, a, b)        srpMakeVec2((v).a, (v).b)
#  define SWZ3(v, a, b, c)     srpMakeVec3((v).a, (v).b, (v).c)
#  define SWZ4(v, a, b, c, d)  srpMakeVec4((v).a, (v).b, (v).c, (v).d)
#endif

#if SRP_INLINE_MATH
#  define SRP_MATH_FN SRP_HD static inline
#  include "srp/detail/vec_impl.h"
#  undef SRP_MATH_FN
#else
SRP_EXTERN_C_BEGIN
vec2 vec2Add(vec2 a, vec2 b);
vec2 vec2Subtract(vec2 a, vec2 b);
float vec2DotProduct(vec2 a, vec2 b);
vec2 vec2MultiplyScalar(vec2 a, float b);
vec2 vec2Normalize(vec2 v);
vec2 vec2Reflect(vec2 i, vec2 n);
vec2 vec2MultiplyVec2(vec2 a, vec2 b);
vec2 vec2Negate(vec2 v);

vec3 vec3Add(vec3 a, vec3 b);
vec3 vec3Subtract(vec3 a, vec3 b);
float vec3DotProduct(vec3 a, vec3 b);
vec3 vec3MultiplyScalar(vec3 a, float b);
vec3 vec3Normalize(vec3 v);
vec3 vec3Reflect(vec3 i, vec3 n);
vec3 vec3MultiplyVec3(vec3 a, vec3 b);
vec3 vec3Negate(vec3 v);

vec4 vec4Add(vec4 a, vec4 b);
vec4 vec4Subtract(vec4 a, vec4 b);
float vec4DotProduct(vec4 a, vec4 b);
vec4 vec4MultiplyScalar(vec4 a, float b);
vec4 vec4Normalize(vec4 v);
vec4 vec4Reflect(vec4 i, vec4 n);
vec4 vec4MultiplyVec4(vec4 a, vec4 b);
vec4 vec4Negate(vec4 v);
SRP_EXTERN_C_END
#endif
