/* Fragment shaders of part 4 of shader_list.h. */
#define SC_DEVICE_PART 4
#include "device_shaders_part.h"
