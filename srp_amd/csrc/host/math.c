/* Exported C definitions of the vec/mat helpers (reference math/vec.c and
 * math/mat.c interfaces) and the host entry of texture sampling.  Built with
 * -ffp-contract=off like every translation unit of the library. */
#define SRP_NO_INLINE_MATH 1
#include <math.h>
#include "srp/vec.h"
#include "srp/mat.h"
#include "srp/texture.h"
#include "srp/srp_hip.h"

#define SRP_MATH_FN
#include "srp/detail/vec_impl.h"
#include "srp/detail/mat_impl.h"
#undef SRP_MATH_FN

void srpTextureGetFilteredColor(const SRPTexture* t, float u, float v, float out[4])
{
	srpTextureGetFilteredColorHost(t, u, v, out);
}
