// Batched multi-shape launches (filled in with the batching layer).
#include <hip/hip_runtime.h>
#include "vg_internal.hpp"
