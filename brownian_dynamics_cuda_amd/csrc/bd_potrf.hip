// bd_potrf.hip -- the blocked Cholesky factor of a caller's matrix in
// device memory: the library's replacement of cusolverDnDpotrf +
// Cholesky_Decomposition_mod (func_cuda.cu:3513-3534, 614-630), with the
// CPU rule of func.cpp:806-836 for a negative pivot (|f|, counted) instead
// of cuSOLVER's failure code.  The factor itself is the one the step uses
// (bd_chol.hip); this file only moves the caller's matrix into the
// factor's padded storage and the factor back out.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/bd_hip.h"
#include "bd_kernels.h"

using namespace bd;

namespace {

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_last_error(buf);
    return code;
}

#define POTRF_CHK(x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return fail(BD_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                        \
    } while (0)

int check_args(int m, const double *D, int ldd, const double *L, int ldl, int mode) {
    if (m < 1) return fail(BD_E_ARG, "m must be >= 1 (got %d)", m);
    if (!D || !L) return fail(BD_E_ARG, "D and L are required");
    if (ldd < m || ldl < m) return fail(BD_E_ARG, "leading dimensions must be >= m");
    if (D == L && ldd != ldl) return fail(BD_E_ARG, "in place needs ldd == ldl");
    if (mode != BD_CHOL_FP64 && mode != BD_CHOL_FP32 && mode != BD_CHOL_BF16)
        return fail(BD_E_ARG, "unsupported chol_mode %d", mode);
    return BD_OK;
}

int split_tiles_of_current_device() {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
    }
    return ncu * 19 / 10;
}

// Work of one factor in the workspace at ws (doubles): the padded matrix
// (mpad x mpad), the inverses of the diagonal factors (one 64x64 tile per
// block column) and the negative-pivot counter.
size_t workspace_doubles(int m) {
    const size_t mpad = ((size_t)m + 63) / 64 * 64;
    return mpad * mpad + (mpad / 64) * 64 * 64 + 32;
}

int run_factor(int m, const double *D, int ldd, double *L, int ldl, int mode, int *neg, hipStream_t q,
               double *ws) {
    CholWork c{};
    c.m = m;
    c.mpad = (m + 63) / 64 * 64;
    c.ld = c.mpad;
    c.A = ws;
    c.Linv = ws + (size_t)c.mpad * c.mpad;
    c.neg = reinterpret_cast<int *>(c.Linv + (size_t)(c.mpad / 64) * 64 * 64);
    c.mode = mode;
    c.split_tiles = split_tiles_of_current_device();
    c.panel_tiles = 0;
    launch_pack(D, ldd, c, q);
    launch_diag(c, 0, true, q);
    launch_cholesky(c, q);
    launch_unpark(c, L, ldl, q);
    POTRF_CHK(hipGetLastError());
    int h = 0;
    POTRF_CHK(hipMemcpyAsync(&h, c.neg, sizeof(int), hipMemcpyDeviceToHost, q));
    POTRF_CHK(hipStreamSynchronize(q));
    if (neg) *neg = h;
    return BD_OK;
}

// Workspace of bd_cholesky_noalloc: a static device array of the library,
// sized for matrices of up to kNoallocMaxM rows (33 block columns; the
// 2082-row tensor of the 100-nucleosome fiber fits): 36.8 MB that every GPU
// the library runs on holds until the process ends.  Host code never uses
// the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, and is kept in g_scratch.
constexpr int kNoallocMaxM = BD_CHOLESKY_NOALLOC_MAX_M;
constexpr size_t kNoallocDoubles =
    (size_t)kNoallocMaxM * kNoallocMaxM + (size_t)(kNoallocMaxM / 64) * 64 * 64 + 32;
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_of_current_device(double **out) {
    int dev = 0;
    POTRF_CHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BD_E_ARG, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(BD_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return BD_OK;
}

}  // namespace

extern "C" {

int bd_cholesky(int m, const double *D, int ldd, double *L, int ldl, int mode, int *neg, void *stream) {
    int r = check_args(m, D, ldd, L, ldl, mode);
    if (r) return r;
    hipStream_t q = static_cast<hipStream_t>(stream);
    double *ws = nullptr;
    POTRF_CHK(hipMalloc((void **)&ws, workspace_doubles(m) * sizeof(double)));
    r = run_factor(m, D, ldd, L, ldl, mode, neg, q, ws);
    hipError_t e = hipFree(ws);
    if (r) return r;
    if (e != hipSuccess) return fail(BD_E_HIP, "hipFree: %s", hipGetErrorString(e));
    return BD_OK;
}

int bd_cholesky_noalloc(int m, const double *D, int ldd, double *L, int ldl, int mode, int *neg, void *stream) {
    int r = check_args(m, D, ldd, L, ldl, mode);
    if (r) return r;
    if (m > kNoallocMaxM)
        return fail(BD_E_CAPACITY, "m=%d exceeds the static workspace of bd_cholesky_noalloc (at most %d rows)", m,
                    kNoallocMaxM);
    double *ws = nullptr;
    r = scratch_of_current_device(&ws);
    if (r) return r;
    return run_factor(m, D, ldd, L, ldl, mode, neg, static_cast<hipStream_t>(stream), ws);
}

}  // extern "C"
