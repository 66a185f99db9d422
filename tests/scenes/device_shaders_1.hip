/* Fragment shaders of part 1 of shader_list.h. */
#define SC_DEVICE_PART 1
#include "device_shaders_part.h"
