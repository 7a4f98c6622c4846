// ORACLE -- test infrastructure only (see vgo.h).
// Navigation restatement (filled in with the navigator).
#include "vgo.h"
