// ORACLE -- test infrastructure only (see vgo.h).
// Remaining shapes of the zoo.  Each is added with its reference citations.
#include "vgo.h"

namespace vgo {

Shape *MakeZooShape(int, const double *, int) { return nullptr; }

}  // namespace vgo
