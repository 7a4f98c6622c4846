// Navigation kernels (filled in with the navigator).
#include <hip/hip_runtime.h>
#include "vg_internal.hpp"
