// PM3 (placeholder until the fine-mesh chain lands): every entry reports
// that the stage is unavailable rather than computing anything.
#include "common.h"

namespace cube2 {
static int unavailable() {
    set_error("PM3 is not available in this build");
    return CUBE2_ERR_STATE;
}
int pm3_force(const int*, const int*, const int*, const int*, const int*, const int64_t*,
              const int16_t*, int64_t, const float* const*, float, float, float, float*, double*) {
    return unavailable();
}
int pm3_debug_job(int, const int*, const int*, const int*, const int*, const int64_t*,
                  const int16_t*, int64_t, const float*, float, float*, float*, float*, float*) {
    return unavailable();
}
int velocity_apply(const int*, const int64_t*, float*, int64_t, float*) { return unavailable(); }
int green_3d(float*, int, float, float, float) { return unavailable(); }
int fft3d(float*, int, int, int) { return unavailable(); }
void pm3_release() {}
int green_cache_get(const float*, size_t, const float**) { return unavailable(); }
void green_cache_release() {}
}  // namespace cube2
