/* Definitions of the vec2/3/4 helpers, written once for the exported C
 * symbols and for the host+device copies HIP shader modules inline.
 * Every expression is evaluated left to right with no fused multiply-add
 * (the build uses -ffp-contract=off), matching the reference's
 * math/vec.c:16-187 results bit for bit.  Requires SRP_MATH_FN. */

SRP_MATH_FN vec2 vec2Add(vec2 a, vec2 b) { vec2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
SRP_MATH_FN vec2 vec2Subtract(vec2 a, vec2 b) { vec2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
SRP_MATH_FN float vec2DotProduct(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
SRP_MATH_FN vec2 vec2MultiplyScalar(vec2 a, float b) { vec2 r; r.x = a.x * b; r.y = a.y * b; return r; }
SRP_MATH_FN vec2 vec2Negate(vec2 v) { vec2 r; r.x = -v.x; r.y = -v.y; return r; }
SRP_MATH_FN vec2 vec2MultiplyVec2(vec2 a, vec2 b) { vec2 r; r.x = a.x * b.x; r.y = a.y * b.y; return r; }
SRP_MATH_FN vec2 vec2Normalize(vec2 v)
{
	vec2 r; r.x = 0.0f; r.y = 0.0f;
	const float len = sqrtf(v.x * v.x + v.y * v.y);
	if (len > 0) { const float s = 1.0f / len; r.x = v.x * s; r.y = v.y * s; }
	return r;
}
SRP_MATH_FN vec2 vec2Reflect(vec2 i, vec2 n)
{
	return vec2Subtract(i, vec2MultiplyScalar(n, 2.f * vec2DotProduct(n, i)));
}

SRP_MATH_FN vec3 vec3Add(vec3 a, vec3 b) { vec3 r; r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; return r; }
SRP_MATH_FN vec3 vec3Subtract(vec3 a, vec3 b) { vec3 r; r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; return r; }
SRP_MATH_FN float vec3DotProduct(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SRP_MATH_FN vec3 vec3MultiplyScalar(vec3 a, float b) { vec3 r; r.x = a.x * b; r.y = a.y * b; r.z = a.z * b; return r; }
SRP_MATH_FN vec3 vec3Negate(vec3 v) { vec3 r; r.x = -v.x; r.y = -v.y; r.z = -v.z; return r; }
SRP_MATH_FN vec3 vec3MultiplyVec3(vec3 a, vec3 b) { vec3 r; r.x = a.x * b.x; r.y = a.y * b.y; r.z = a.z * b.z; return r; }
SRP_MATH_FN vec3 vec3Normalize(vec3 v)
{
	vec3 r; r.x = 0.0f; r.y = 0.0f; r.z = 0.0f;
	const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
	if (len > 0) { const float s = 1 / len; r.x = v.x * s; r.y = v.y * s; r.z = v.z * s; }
	return r;
}
SRP_MATH_FN vec3 vec3Reflect(vec3 i, vec3 n)
{
	return vec3Subtract(i, vec3MultiplyScalar(n, 2.f * vec3DotProduct(n, i)));
}

SRP_MATH_FN vec4 vec4Add(vec4 a, vec4 b) { vec4 r; r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; r.w = a.w + b.w; return r; }
SRP_MATH_FN vec4 vec4Subtract(vec4 a, vec4 b) { vec4 r; r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; r.w = a.w - b.w; return r; }
SRP_MATH_FN float vec4DotProduct(vec4 a, vec4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
SRP_MATH_FN vec4 vec4MultiplyScalar(vec4 a, float b) { vec4 r; r.x = a.x * b; r.y = a.y * b; r.z = a.z * b; r.w = a.w * b; return r; }
SRP_MATH_FN vec4 vec4Negate(vec4 v) { vec4 r; r.x = -v.x; r.y = -v.y; r.z = -v.z; r.w = -v.w; return r; }
SRP_MATH_FN vec4 vec4MultiplyVec4(vec4 a, vec4 b) { vec4 r; r.x = a.x * b.x; r.y = a.y * b.y; r.z = a.z * b.z; r.w = a.w * b.w; return r; }
SRP_MATH_FN vec4 vec4Normalize(vec4 v)
{
	vec4 r; r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 0.0f;
	const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w);
	if (len > 0) { const float s = 1.0f / len; r.x = v.x * s; r.y = v.y * s; r.z = v.z * s; r.w = v.w * s; }
	return r;
}
SRP_MATH_FN vec4 vec4Reflect(vec4 i, vec4 n)
{
	return vec4Subtract(i, vec4MultiplyScalar(n, 2.f * vec4DotProduct(n, i)));
}
