// Serving loop (filled in with the serving layer).
#include "vg_internal.hpp"
