/* Device instances of the scene shaders (GPU build of the scene module):
 * the file that includes this defines SC_DEVICE_PART and gets the kernels
 * of that part of shader_list.h, so the parts compile in parallel. */
#include <srp/srp_device.h>
#include "shaders.h"

#if SC_DEVICE_PART == 0
#define SC_PART_0(x) x
#else
#define SC_PART_0(x)
#endif
#if SC_DEVICE_PART == 1
#define SC_PART_1(x) x
#else
#define SC_PART_1(x)
#endif
#if SC_DEVICE_PART == 2
#define SC_PART_2(x) x
#else
#define SC_PART_2(x)
#endif
#if SC_DEVICE_PART == 3
#define SC_PART_3(x) x
#else
#define SC_PART_3(x)
#endif
#if SC_DEVICE_PART == 4
#define SC_PART_4(x) x
#else
#define SC_PART_4(x)
#endif

#define HOST_VS(name, fn, usize, part) SC_PART_##part(SRP_HIP_VERTEX_SHADER(name, fn, usize))
#define HOST_FS(name, fn, usize, part) SC_PART_##part(SRP_HIP_FRAGMENT_SHADER(name, fn, usize))
#include "shader_list.h"
