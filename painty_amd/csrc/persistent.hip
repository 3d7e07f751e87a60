// Whole stamp lists in one persistent launch.
#include "internal.hpp"
#include "persistent.hpp"

namespace pnt {

struct PersistentState {};

void run_persistent(pnt_brush_s*, pnt_canvas_s*, const pnt_stamp*, int64_t) {
  throw Error(PNT_ESTATE, "PNT_EXEC_PERSISTENT is not available in this build");
}

void persistent_release(PersistentState*& p) {
  delete p;
  p = nullptr;
}

}  // namespace pnt
