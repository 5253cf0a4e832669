// Green's table cache files (initialize.f90:102-175).
//
// The reference writes each table with access='stream': the raw real(4)
// values of Gk(n/2+1, n, n) in Fortran (column-major) order, no record
// markers, which is the byte image of the device table d_G[n][n][n/2+1].
// A read fills the table from the first n*n*(n/2+1) floats of the file and
// fails on a shorter file, as the Fortran read does at end of file.
#include <cstdio>
#include <vector>

#include "engine.h"

namespace cube2 {

static size_t green_elems(int n) { return (size_t)n * n * (size_t)(n / 2 + 1); }

int green_file_write(const float* d_G, int n, const char* path) {
    if (!d_G || !path || n < 2 || n % 2) {
        set_error("green file write: bad arguments (n=%d)", n);
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    const size_t m = green_elems(n);
    std::vector<float> h(m);
    hipStream_t s = ctx().stream;
    C2_HIP(hipMemcpyAsync(h.data(), d_G, m * sizeof(float), hipMemcpyDeviceToHost, s));
    C2_HIP(hipStreamSynchronize(s));
    std::FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error("green file write: cannot open %s", path);
        return CUBE2_ERR_ARG;
    }
    const size_t w = std::fwrite(h.data(), sizeof(float), m, f);
    const int c = std::fclose(f);
    if (w != m || c != 0) {
        set_error("green file write: short write to %s", path);
        return CUBE2_ERR_ARG;
    }
    return CUBE2_OK;
}

int green_file_read(float* d_G, int n, const char* path) {
    if (!d_G || !path || n < 2 || n % 2) {
        set_error("green file read: bad arguments (n=%d)", n);
        return CUBE2_ERR_ARG;
    }
    const size_t m = green_elems(n);
    std::FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error("green file read: cannot open %s", path);
        return CUBE2_ERR_ARG;
    }
    std::vector<float> h(m);
    const size_t r = std::fread(h.data(), sizeof(float), m, f);
    std::fclose(f);
    if (r != m) {
        set_error("green file read: %s holds %zu of the %zu floats of a %d^2 x %d table", path, r,
                  m, n, n / 2 + 1);
        return CUBE2_ERR_ARG;
    }
    C2_TRY(ensure_init());
    hipStream_t s = ctx().stream;
    C2_HIP(hipMemcpyAsync(d_G, h.data(), m * sizeof(float), hipMemcpyHostToDevice, s));
    C2_HIP(hipStreamSynchronize(s));
    return CUBE2_OK;
}

}  // namespace cube2
