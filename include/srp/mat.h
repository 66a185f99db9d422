/* mat4 (row-major 4x4).  Interface of reference include/srp/mat.h:16-93. */
#pragma once

#include "srp/vec.h"
#include "srp/detail/qualifiers.h"

typedef struct mat4 { float data[4][4]; } mat4;

#if SRP_INLINE_MATH
#  define SRP_MATH_FN SRP_HD static inline
#  include "srp/detail/mat_impl.h"
#  undef SRP_MATH_FN
#else
SRP_EXTERN_C_BEGIN
vec4 mat4MultiplyVec4(const mat4* SRP_RESTRICT a, vec4 b);
mat4 mat4MultiplyMat4(const mat4* SRP_RESTRICT a, const mat4* SRP_RESTRICT b);
mat4 mat4ConstructIdentity(void);
mat4 mat4ConstructScale(float x, float y, float z);
mat4 mat4ConstructTranslate(float x, float y, float z);
/** Angles in radians about X, Y, Z (entries evaluated in double precision). */
mat4 mat4ConstructRotate(float x, float y, float z);
/** Translate * (Rotate * Scale). */
mat4 mat4ConstructTRS(
	float transX, float transY, float transZ,
	float rotationX, float rotationY, float rotationZ,
	float scaleX, float scaleY, float scaleZ);
/** TRS with negated translation and rotation. */
mat4 mat4ConstructView(
	float cameraX, float cameraY, float cameraZ,
	float rotationX, float rotationY, float rotationZ,
	float scaleX, float scaleY, float scaleZ);
mat4 mat4ConstructOrthogonalProjection(
	float x_min, float x_max, float y_min, float y_max, float z_min, float z_max);
mat4 mat4ConstructPerspectiveProjection(
	float x_min_near, float x_max_near, float y_min_near, float y_max_near,
	float z_near, float z_far);
SRP_EXTERN_C_END
#endif
