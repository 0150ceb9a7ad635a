// v2d_force_coeffs_noalloc: the force integral of v2d_force_coeffs with no
// device allocation.  The integral (k_forces) reads only the wall layers:
// row j = 1 of beta, gama, eta_x, eta_y, x, y and rows j = 1..3 of omega.  The
// allocating entry uploads whole planes into buffers it allocates per call;
// this one stages the wall rows, and receives the ten coefficients, in a
// file-scope device array of fixed size (kScratchDoubles doubles, 1.7 MB),
// which every GPU the library runs on holds until the process ends.
//
// Scratch layout for a mesh of imax wall columns: ten metric planes of stride
// imax (planes 2, 3, 6, 7, 8, 9 filled, the others unread), three omega rows,
// then the ten outputs.  Every call overwrites all of what the kernel reads,
// so nothing is cleared; calls that overlap on one GPU (two host threads, one
// device) race on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/vorti2d.h"
#include "v2d_internal.h"

namespace v2d {
namespace {

constexpr int kMaxImax = V2D_NOALLOC_MAX_IMAX;
constexpr int64_t kScratchDoubles = 13 * (int64_t)kMaxImax + 16;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch`, looked up once per
// device (with that device current) and kept.
int scratch_ptr(double **out)
{
    int dev = 0;
    V2D_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("force coefficients (no-alloc): device index " + std::to_string(dev) +
                  " beyond the " + std::to_string(kMaxDevices) + " the scratch table holds");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("force coefficients (no-alloc): hipGetSymbolAddress of the scratch "
                                  "array failed: ") + hipGetErrorString(e));
            return -1;
        }
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace
}  // namespace v2d

using namespace v2d;

extern "C" int v2d_force_coeffs_noalloc(int imax, int jmax, double re, double dksi, double deta,
                                        const double *beta, const double *gama,
                                        const double *detadx, const double *detady,
                                        const double *xphys, const double *yphys,
                                        const double *ome, double ref_length, double x0, double y0,
                                        double *out)
{
    if (imax < 4 || jmax < 4) {
        set_error("force coefficients (no-alloc): imax and jmax must be >= 4");
        return -1;
    }
    if (imax > kMaxImax) {
        set_error("force coefficients (no-alloc): imax " + std::to_string(imax) + " exceeds the " +
                  std::to_string(kMaxImax) + " wall columns its fixed scratch holds "
                  "(use v2d_force_coeffs)");
        return -1;
    }
    double *s = nullptr;
    if (scratch_ptr(&s)) return -1;
    const int64_t w = imax;
    double *m = s;                 // 10 planes of stride imax
    double *ow = s + 10 * w;       // omega rows j = 1, 2, 3
    double *dout = s + 13 * w;     // the ten coefficients
    const double *planes[10] = {nullptr, nullptr, beta, gama, nullptr, nullptr, detadx, detady, xphys, yphys};
    for (int q = 0; q < 10; ++q)
        if (planes[q]) V2D_HIP(hipMemcpy(m + q * w, planes[q], sizeof(double) * w, hipMemcpyHostToDevice));
    V2D_HIP(hipMemcpy(ow, ome, sizeof(double) * 3 * w, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(1024), 0, nullptr, imax, re, dksi, deta, m, w, ow,
                       ref_length, x0, y0, dout);
    V2D_LAUNCH();
    V2D_HIP(hipMemcpy(out, dout, sizeof(double) * 10, hipMemcpyDeviceToHost));
    return 0;
}
