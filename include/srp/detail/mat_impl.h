/* Definitions of the mat4 helpers (exported C symbols and HIP inline copies).
 * Products are summed left to right without contraction, and the rotation
 * entries are formed in double precision from float angles before rounding
 * to float, which reproduces reference math/mat.c:16-215.  On the device
 * side sin/cos come from the device math library, so matrices should be
 * built on the host (as every reference program does). Requires SRP_MATH_FN. */
#include <math.h>

SRP_MATH_FN vec4 mat4MultiplyVec4(const mat4* SRP_RESTRICT m, vec4 v)
{
	vec4 r;
	r.x = m->data[0][0] * v.x + m->data[0][1] * v.y + m->data[0][2] * v.z + m->data[0][3] * v.w;
	r.y = m->data[1][0] * v.x + m->data[1][1] * v.y + m->data[1][2] * v.z + m->data[1][3] * v.w;
	r.z = m->data[2][0] * v.x + m->data[2][1] * v.y + m->data[2][2] * v.z + m->data[2][3] * v.w;
	r.w = m->data[3][0] * v.x + m->data[3][1] * v.y + m->data[3][2] * v.z + m->data[3][3] * v.w;
	return r;
}

SRP_MATH_FN mat4 mat4MultiplyMat4(const mat4* SRP_RESTRICT a, const mat4* SRP_RESTRICT b)
{
	mat4 r;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			r.data[i][j] = a->data[i][0] * b->data[0][j] + a->data[i][1] * b->data[1][j]
			             + a->data[i][2] * b->data[2][j] + a->data[i][3] * b->data[3][j];
	return r;
}

SRP_HD static inline mat4 srpMat4Diag_(float a, float b, float c, float d)
{
	mat4 r;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			r.data[i][j] = 0.0f;
	r.data[0][0] = a; r.data[1][1] = b; r.data[2][2] = c; r.data[3][3] = d;
	return r;
}

SRP_MATH_FN mat4 mat4ConstructIdentity(void) { return srpMat4Diag_(1, 1, 1, 1); }
SRP_MATH_FN mat4 mat4ConstructScale(float x, float y, float z) { return srpMat4Diag_(x, y, z, 1); }

SRP_MATH_FN mat4 mat4ConstructTranslate(float x, float y, float z)
{
	mat4 r = srpMat4Diag_(1, 1, 1, 1);
	r.data[0][3] = x; r.data[1][3] = y; r.data[2][3] = z;
	return r;
}

SRP_MATH_FN mat4 mat4ConstructRotate(float x, float y, float z)
{
	const double sx = sin((double) x), cx = cos((double) x);
	const double sy = sin((double) y), cy = cos((double) y);
	const double sz = sin((double) z), cz = cos((double) z);
	mat4 r = srpMat4Diag_(0, 0, 0, 1);
	r.data[0][0] = cy * cz;
	r.data[0][1] = sx * sy * cz - cx * sz;
	r.data[0][2] = cx * sy * cz + sx * sz;
	r.data[1][0] = cy * sz;
	r.data[1][1] = sx * sy * sz + cx * cz;
	r.data[1][2] = cx * sy * sz - sx * cz;
	r.data[2][0] = -sy;
	r.data[2][1] = sx * cy;
	r.data[2][2] = cx * cy;
	return r;
}

SRP_MATH_FN mat4 mat4ConstructTRS(
	float transX, float transY, float transZ,
	float rotationX, float rotationY, float rotationZ,
	float scaleX, float scaleY, float scaleZ)
{
	const mat4 t = mat4ConstructTranslate(transX, transY, transZ);
	const mat4 rot = mat4ConstructRotate(rotationX, rotationY, rotationZ);
	const mat4 s = mat4ConstructScale(scaleX, scaleY, scaleZ);
	const mat4 rs = mat4MultiplyMat4(&rot, &s);
	return mat4MultiplyMat4(&t, &rs);
}

SRP_MATH_FN mat4 mat4ConstructView(
	float cameraX, float cameraY, float cameraZ,
	float rotationX, float rotationY, float rotationZ,
	float scaleX, float scaleY, float scaleZ)
{
	return mat4ConstructTRS(-cameraX, -cameraY, -cameraZ,
	                        -rotationX, -rotationY, -rotationZ,
	                        scaleX, scaleY, scaleZ);
}

SRP_MATH_FN mat4 mat4ConstructOrthogonalProjection(
	float x_min, float x_max, float y_min, float y_max, float z_min, float z_max)
{
	mat4 r = srpMat4Diag_(2 / (x_max - x_min), 2 / (y_max - y_min), 2 / (z_max - z_min), 1);
	r.data[0][3] = -(x_max + x_min) / (x_max - x_min);
	r.data[1][3] = -(y_max + y_min) / (y_max - y_min);
	r.data[2][3] = -(z_max + z_min) / (z_max - z_min);
	return r;
}

SRP_MATH_FN mat4 mat4ConstructPerspectiveProjection(
	float x_min_near, float x_max_near, float y_min_near, float y_max_near,
	float z_near, float z_far)
{
	/* Perspective part (maps the frustum to a box, w := z) followed by the
	 * orthographic map of that box to the unit cube. */
	mat4 p = srpMat4Diag_(z_near, z_near, z_near + z_far, 0);
	p.data[2][3] = -z_near * z_far;
	p.data[3][2] = 1;
	const mat4 o = mat4ConstructOrthogonalProjection(
		x_min_near, x_max_near, y_min_near, y_max_near, z_near, z_far);
	return mat4MultiplyMat4(&o, &p);
}
