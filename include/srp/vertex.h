/* Vertex and varying types.  Interface of reference include/srp/vertex.h:17-46. */
#pragma once

#include <stddef.h>
#include "srp/type.h"

/** User vertex layout (opaque to the library; stride given at upload). */
typedef struct SRPVertex SRPVertex;

/** Per-vertex outputs of the vertex shader (raw bytes described by SRPVaryingInfo). */
typedef struct SRPVarying SRPVarying;

typedef enum SRPInterpolationMode
{
	SRP_INTERPOLATION_MODE_PERSPECTIVE,  /**< divide by interpolated 1/w */
	SRP_INTERPOLATION_MODE_AFFINE,       /**< screen-space linear */
	SRP_INTERPOLATION_MODE_FLAT          /**< value of the provoking vertex */
} SRPInterpolationMode;

/** One varying: `nItems` consecutive elements of `type`. */
typedef struct
{
	size_t nItems;
	SRPType type;
	SRPInterpolationMode interpolationMode;
} SRPVaryingInfo;

/** Interpolated varyings as seen by the fragment shader. */
typedef struct SRPInterpolated SRPInterpolated;
