/* Device instances of the scene shaders (GPU build of the scene module). */
#include <srp/srp_device.h>
#include "shaders.h"

#define HOST_VS(name, fn, usize) SRP_HIP_VERTEX_SHADER(name, fn, usize)
#define HOST_FS(name, fn, usize) SRP_HIP_FRAGMENT_SHADER(name, fn, usize)
#include "shader_list.h"
