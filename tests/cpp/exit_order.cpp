// Regression check (no device needed): a caller's static object that calls
// into the library while the process exits -- after the main thread's
// thread-local objects have been destroyed -- must find the library's
// per-thread error text still valid.  (It used to be a std::string: by
// then its buffer was freed, and the next call's bookkeeping wrote into the
// freed chunk, corrupting the allocator's free lists.)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "painty_hip.h"

namespace {
const char kMsg[] = "Palette size does not match underlying size.";
struct AtExit {
  ~AtExit() {
    const bool intact = std::strcmp(pnt_last_error(), kMsg) == 0;
    pnt_ctx_destroy(nullptr);  // a no-op call that runs the error bookkeeping
    const bool cleared = pnt_last_error()[0] == '\0';
    std::printf("%s\n", intact && cleared ? "exit ok" : "exit BROKEN");
    std::fflush(stdout);
    if (!(intact && cleared)) std::_Exit(3);
  }
};
}  // namespace

int main() {
  static AtExit at_exit;
  double K[3], S[3];
  const double w[2] = {0.5, 0.5}, bk[3] = {1, 1, 1}, bs[3] = {1, 1, 1};
  // one palette entry, two weights: the reference's size-mismatch error
  if (pnt_mix_single(1, bk, bs, 2, w, K, S) != PNT_EINVAL) return 1;
  if (std::strcmp(pnt_last_error(), kMsg) != 0) return 2;
  return 0;
}
