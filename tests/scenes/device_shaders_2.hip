/* Fragment shaders of part 2 of shader_list.h. */
#define SC_DEVICE_PART 2
#include "device_shaders_part.h"
