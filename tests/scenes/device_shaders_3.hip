/* Fragment shaders of part 3 of shader_list.h. */
#define SC_DEVICE_PART 3
#include "device_shaders_part.h"
