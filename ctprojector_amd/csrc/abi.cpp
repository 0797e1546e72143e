// C ABI of libprojector.so.  The exported symbols, argument lists and
// semantics are those of the reference's libprojector.so (declared in
// include/ctprojector.h with the reference file:line each replaces):
//   c*     host pointers: staged through device buffers allocated per call;
//   cupy*  device pointers owned by the caller (any HIP device memory, e.g.
//          torch tensors); outputs must be zero-initialised where the reference
//          accumulates.
// Every call runs on the current device's null stream and returns when the
// results are complete.  Failures never abort: the function prints
// "<name>() failed: <reason>" to stderr like the reference and returns a
// non-zero hipError_t code (0 = success); ctpGetLastErrorMessage() returns the
// text of the calling thread's last failure.
#include "common.h"
#include "ops.h"
#include "runtime.h"

#include <cstdio>
#include <cstring>
#include <iostream>
#include <algorithm>
#include <limits>
#include <string>
#include <vector>

using namespace ctp;

namespace {

// set by every helper that touches the device during a call; calls that only
// validate their arguments (or have nothing to do) never initialise HIP
thread_local bool t_device_used = false;

template <typename F>
int guarded(const char* name, F&& f) {
    t_device_used = false;
    try {
        f();
        if (t_device_used) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) throw Error((int)e, std::string("kernel failure: ") + hipGetErrorString(e));
            timing_collect();
        }
        return 0;
    } catch (const Error& e) {
        std::string m = std::string(name) + "() failed: " + e.what();
        set_last_error(m);
        std::cerr << m << std::endl;
        return e.code ? e.code : (int)hipErrorUnknown;
    } catch (const std::exception& e) {
        std::string m = std::string(name) + "() failed: " + e.what();
        set_last_error(m);
        std::cerr << m << std::endl;
        return (int)hipErrorUnknown;
    } catch (...) {
        std::string m = std::string(name) + "() failed: unknown exception";
        set_last_error(m);
        std::cerr << m << std::endl;
        return (int)hipErrorUnknown;
    }
}

constexpr size_t kIntMax = (size_t)std::numeric_limits<int>::max();

void need(bool ok, const std::string& msg) {
    if (!ok) fail(msg);
}

Grid make_grid(size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz) {
    need(nx <= kIntMax && ny <= kIntMax && nz <= kIntMax, "image dimension too large");
    need(nx * ny <= kIntMax && nx * ny * nz <= kIntMax, "image of more than 2^31-1 voxels per batch item");
    need(ny <= 65535 && nz <= 65535, "ny or nz above 65535");
    need(dx != 0 && dy != 0 && dz != 0, "zero voxel size");
    return Grid{(int)nx, (int)ny, (int)nz, dx, dy, dz, cx, cy, cz};
}

Det make_det(size_t nu, size_t nv, size_t nview, float du, float dv, float off_u, float off_v) {
    need(nu <= kIntMax && nv <= kIntMax && nview <= kIntMax, "detector dimension too large");
    need(nview <= 65535, "more than 65535 views per call");
    need(nv <= 65535 * 16, "more than 1048560 detector rows");
    need(nu * nv <= kIntMax, "detector of more than 2^31-1 cells");
    need(du != 0 && dv != 0, "zero detector pitch");
    return Det{(int)nu, (int)nv, du, dv, off_u, off_v};
}

bool empty(size_t B, size_t a, size_t b, size_t c, size_t d, size_t e, size_t f) {
    return B == 0 || a == 0 || b == 0 || c == 0 || d == 0 || e == 0 || f == 0;
}

void need_ptr(const void* p, const char* what) {
    if (!p) fail(std::string("null pointer: ") + what);
}

// Stream of the calling thread's library calls (ctpSetStream); the null stream
// by default.  Every entry point enqueues its work there and synchronises only
// that stream, so threads with their own streams overlap on the device.
thread_local hipStream_t t_stream = nullptr;
hipStream_t cur_stream() { return t_stream; }

// first device touch of a call: clear errors left over from earlier work
void touch() {
    if (!t_device_used) {
        t_device_used = true;
        (void)hipGetLastError();
    }
}

void up(void* dst, const void* src, size_t bytes) {
    touch();
    if (!bytes) return;
    CTP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cur_stream()));
    CTP_CHECK(hipStreamSynchronize(cur_stream()));
}
void down(void* dst, const void* src, size_t bytes) {
    touch();
    if (!bytes) return;
    CTP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, cur_stream()));
    CTP_CHECK(hipStreamSynchronize(cur_stream()));
}
std::vector<float> host_copy(const float* dev, size_t n) {
    std::vector<float> h(n);
    down(h.data(), dev, n * sizeof(float));
    return h;
}
void zero(void* dev, size_t bytes) {
    touch();
    if (bytes) CTP_CHECK(hipMemsetAsync(dev, 0, bytes, cur_stream()));
}
void sync() {
    touch();
    CTP_CHECK(hipStreamSynchronize(cur_stream()));
}

}  // namespace

#define CTP_EXPORT extern "C" __attribute__((visibility("default")))

CTP_EXPORT int SetDevice(int device) {
    (void)hipGetLastError();
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return (int)e;
    if (device < 0 || device >= n) return (int)hipErrorInvalidDevice;
    return (int)hipSetDevice(device);
}

CTP_EXPORT const char* ctpGetLastErrorMessage(void) { return last_error(); }

// Sets the HIP stream (a hipStream_t; NULL = the null stream) on which the
// calling thread's subsequent library calls run.
CTP_EXPORT int ctpSetStream(void* stream) {
    t_stream = static_cast<hipStream_t>(stream);
    return 0;
}

// Kernel timing: when enabled, every kernel launch of the library is bracketed
// by HIP events on its stream and the elapsed times are accumulated per kernel
// name (for measurement; adds two event records per launch).
CTP_EXPORT int ctpSetKernelTiming(int enable) {
    timing_enable(enable != 0);
    return 0;
}
CTP_EXPORT int ctpResetKernelTimes(void) {
    timing_reset();
    return 0;
}
CTP_EXPORT int ctpGetKernelTime(const char* name, double* total_ms, long* count) {
    return timing_get(name, total_ms, count) ? 0 : (int)hipErrorNotFound;
}
// Writes "name total_ms count" lines into buf (NUL-terminated, truncated to n).
CTP_EXPORT int ctpKernelTimeReport(char* buf, size_t n) {
    std::string r = timing_report();
    if (buf && n) {
        size_t k = std::min(n - 1, r.size());
        std::memcpy(buf, r.data(), k);
        buf[k] = 0;
    }
    return (int)r.size();
}

// Launch-shape knobs (see runtime.h); returns hipErrorInvalidValue for an
// unknown key.
CTP_EXPORT int ctpSetTuning(const char* key, int value) {
    if (!key || !tuning_set(key, value)) return (int)hipErrorInvalidValue;
    return 0;
}
CTP_EXPORT int ctpGetTuning(const char* key) { return key ? tuning_get(key, 0) : 0; }

CTP_EXPORT int ctpReleaseWorkspace(void) {
    return guarded("ctpReleaseWorkspace", [&] {
        release_all_workspaces();
        release_dd_call_cache();
    });
}

// ===========================================================================
// Siddon cone beam, arbitrary trajectory (cuda/siddonCone.cu:137-424)
// ===========================================================================
#define CONE_ARGS                                                                                                   \
    const float *detCenter, const float *detU, const float *detV, const float *src, size_t nBatches, size_t nx,   \
        size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu, size_t nv,     \
        size_t nview, float du, float dv, float off_u, float off_v

static void cone_common(size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx,
                        float cy, float cz, size_t nu, size_t nv, size_t nview, float du, float dv, float off_u,
                        float off_v, Grid& g, Det& d) {
    g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
    d = make_det(nu, nv, nview, du, dv, off_u, off_v);
    need(nBatches <= kIntMax, "batch too large");
}

CTP_EXPORT int cupySiddonConeProjectionArbitrary(float* prj, const float* img, CONE_ARGS) {
    return guarded("cupySiddonConeProjectionArbitrary", [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g; Det d;
        cone_common(nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, g, d);
        need_ptr(prj, "prj"); need_ptr(img, "img"); need_ptr(detCenter, "detCenter"); need_ptr(detU, "detU");
        need_ptr(detV, "detV"); need_ptr(src, "src");
        touch();
        siddon_cone_fp(prj, img, ConeGeomDev{detCenter, detU, detV, src}, (int)nBatches, g, d, (int)nview, cur_stream());
        sync();
    });
}

CTP_EXPORT int cSiddonConeProjectionArbitrary(float* prj, const float* img, CONE_ARGS) {
    return guarded("cSiddonConeProjectionArbitrary", [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g; Det d;
        cone_common(nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, g, d);
        need_ptr(prj, "prj"); need_ptr(img, "img"); need_ptr(detCenter, "detCenter"); need_ptr(detU, "detU");
        need_ptr(detV, "detV"); need_ptr(src, "src");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview, ng = 3 * nview;
        touch();
        DeviceBuffer di(nimg * 4), dp(nprj * 4), dg(4 * ng * 4);
        float* geo = dg.as<float>();
        up(di.p, img, nimg * 4);
        up(geo, detCenter, ng * 4); up(geo + ng, detU, ng * 4); up(geo + 2 * ng, detV, ng * 4); up(geo + 3 * ng, src, ng * 4);
        zero(dp.p, nprj * 4);   // the host API returns A img (cuda/siddonCone.cu:209)
        siddon_cone_fp(dp.as<float>(), di.as<float>(), ConeGeomDev{geo, geo + ng, geo + 2 * ng, geo + 3 * ng},
                       (int)nBatches, g, d, (int)nview, cur_stream());
        sync();
        down(prj, dp.p, nprj * 4);
    });
}

CTP_EXPORT int cupySiddonConeBackprojectionArbitrary(float* img, const float* prj, CONE_ARGS) {
    return guarded("cupySiddonConeBackprojectionArbitrary", [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g; Det d;
        cone_common(nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, g, d);
        need_ptr(prj, "prj"); need_ptr(img, "img"); need_ptr(detCenter, "detCenter"); need_ptr(detU, "detU");
        need_ptr(detV, "detV"); need_ptr(src, "src");
        const size_t ng = 3 * nview;
        std::vector<float> hc = host_copy(detCenter, ng), hu = host_copy(detU, ng), hv = host_copy(detV, ng),
                           hs = host_copy(src, ng);
        siddon_cone_bp(img, prj, ConeGeomDev{detCenter, detU, detV, src},
                       ConeGeomHost{hc.data(), hu.data(), hv.data(), hs.data()}, (int)nBatches, g, d, (int)nview,
                       cur_stream());
        sync();
    });
}

CTP_EXPORT int cSiddonConeBackprojectionArbitrary(float* img, const float* prj, CONE_ARGS) {
    return guarded("cSiddonConeBackprojectionArbitrary", [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g; Det d;
        cone_common(nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, g, d);
        need_ptr(prj, "prj"); need_ptr(img, "img"); need_ptr(detCenter, "detCenter"); need_ptr(detU, "detU");
        need_ptr(detV, "detV"); need_ptr(src, "src");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview, ng = 3 * nview;
        touch();
        DeviceBuffer di(nimg * 4), dp(nprj * 4), dg(4 * ng * 4);
        float* geo = dg.as<float>();
        up(dp.p, prj, nprj * 4);
        up(geo, detCenter, ng * 4); up(geo + ng, detU, ng * 4); up(geo + 2 * ng, detV, ng * 4); up(geo + 3 * ng, src, ng * 4);
        zero(di.p, nimg * 4);
        siddon_cone_bp(di.as<float>(), dp.as<float>(), ConeGeomDev{geo, geo + ng, geo + 2 * ng, geo + 3 * ng},
                       ConeGeomHost{detCenter, detU, detV, src}, (int)nBatches, g, d, (int)nview, cur_stream());
        sync();
        down(img, di.p, nimg * 4);
    });
}

// ===========================================================================
// Distance driven fan / parallel (cuda/distanceDriven2D.cu:109-634)
// ===========================================================================
#define DD_GRID_ARGS                                                                                          \
    size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, \
        size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v

template <bool FP, bool DEV>
static int dd_entry(const char* name, float* out, const float* in, const float* deg, size_t nBatches, size_t nx,
                    size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                    size_t nv, size_t nview, float da, float dv, float off_a, float off_v, float dsd, float dso,
                    int type, bool fan) {
    return guarded(name, [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need(nBatches <= kIntMax, "batch too large");
        need_ptr(out, "out"); need_ptr(in, "in"); need_ptr(deg, "angles");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview;
        if (DEV) {
            std::vector<float> hdeg = host_copy(deg, nview);
            if (FP) dd2d_fp(out, in, hdeg.data(), (int)nBatches, g, d, (int)nview, dsd, dso, type, fan, cur_stream());
            else dd2d_bp(out, in, hdeg.data(), (int)nBatches, g, d, (int)nview, dsd, dso, type, fan, cur_stream());
            sync();
        } else {
            const size_t nin = FP ? nimg : nprj, nout = FP ? nprj : nimg;
            touch();
            DeviceBuffer di(nin * 4), dout(nout * 4);
            up(di.p, in, nin * 4);
            zero(dout.p, nout * 4);
            if (FP) dd2d_fp(dout.as<float>(), di.as<float>(), deg, (int)nBatches, g, d, (int)nview, dsd, dso, type, fan, cur_stream());
            else dd2d_bp(dout.as<float>(), di.as<float>(), deg, (int)nBatches, g, d, (int)nview, dsd, dso, type, fan, cur_stream());
            sync();
            down(out, dout.p, nout * 4);
        }
    });
}

CTP_EXPORT int cupyDistanceDrivenFanProjection(float* prj, const float* img, const float* deg, DD_GRID_ARGS, float dsd,
                                               float dso, int typeProjector) {
    return dd_entry<true, true>("cupyDistanceDrivenFanProjection", prj, img, deg, nBatches, nx, ny, nz, dx, dy, dz, cx,
                                cy, cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso, typeProjector, true);
}
// No counterpart in the reference: cupyDistanceDrivenFanProjection for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  The scratch is a fixed
// 80 MiB device array per GPU; calls needing more fail with an error; calls
// that overlap on one GPU race on it.
CTP_EXPORT int ctpDistanceDrivenFanProjectionNoAlloc(float* prj, const float* img, const float* deg, DD_GRID_ARGS,
                                                     float dsd, float dso, int typeProjector) {
    return guarded("ctpDistanceDrivenFanProjectionNoAlloc", [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need(nBatches <= kIntMax, "batch too large");
        need_ptr(prj, "prj"); need_ptr(img, "img"); need_ptr(deg, "angles");
        std::vector<float> hdeg = host_copy(deg, nview);
        dd2d_fp_noalloc(prj, img, hdeg.data(), (int)nBatches, g, d, (int)nview, dsd, dso, typeProjector, true,
                        cur_stream());
        sync();
    });
}
CTP_EXPORT int cDistanceDrivenFanProjection(float* prj, const float* img, const float* deg, DD_GRID_ARGS, float dsd,
                                            float dso, int typeProjector) {
    return dd_entry<true, false>("cDistanceDrivenFanProjection", prj, img, deg, nBatches, nx, ny, nz, dx, dy, dz, cx,
                                 cy, cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso, typeProjector, true);
}
CTP_EXPORT int cupyDistanceDrivenParallelProjection(float* prj, const float* img, const float* deg, DD_GRID_ARGS,
                                                    int typeProjector) {
    return dd_entry<true, true>("cupyDistanceDrivenParallelProjection", prj, img, deg, nBatches, nx, ny, nz, dx, dy, dz,
                                cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, typeProjector, false);
}
CTP_EXPORT int cDistanceDrivenParallelProjection(float* prj, const float* img, const float* deg, DD_GRID_ARGS,
                                                 int typeProjector) {
    return dd_entry<true, false>("cDistanceDrivenParallelProjection", prj, img, deg, nBatches, nx, ny, nz, dx, dy, dz,
                                 cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, typeProjector, false);
}
CTP_EXPORT int cupyDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg, DD_GRID_ARGS,
                                                   float dsd, float dso, int typeProjector) {
    return dd_entry<false, true>("cupyDistanceDrivenFanBackprojection", img, prj, deg, nBatches, nx, ny, nz, dx, dy, dz,
                                 cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso, typeProjector, true);
}
CTP_EXPORT int cDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg, DD_GRID_ARGS, float dsd,
                                                float dso, int typeProjector) {
    return dd_entry<false, false>("cDistanceDrivenFanBackprojection", img, prj, deg, nBatches, nx, ny, nz, dx, dy, dz,
                                  cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso, typeProjector, true);
}
CTP_EXPORT int cupyDistanceDrivenParallelBackprojection(float* img, const float* prj, const float* deg, DD_GRID_ARGS,
                                                        int typeProjector) {
    return dd_entry<false, true>("cupyDistanceDrivenParallelBackprojection", img, prj, deg, nBatches, nx, ny, nz, dx,
                                 dy, dz, cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, typeProjector,
                                 false);
}
CTP_EXPORT int cDistanceDrivenParallelBackprojection(float* img, const float* prj, const float* deg, DD_GRID_ARGS,
                                                     int typeProjector) {
    return dd_entry<false, false>("cDistanceDrivenParallelBackprojection", img, prj, deg, nBatches, nx, ny, nz, dx, dy,
                                  dz, cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, typeProjector,
                                  false);
}

// img += A^T prj, the exact transpose of the distance-driven forward projector
// (device pointers; no counterpart in the reference): for gradients.
static int dd_adjoint_entry(const char* name, float* img, const float* prj, const float* deg, size_t nBatches,
                            size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                            size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v, float dsd,
                            float dso, int type, bool fan) {
    return guarded(name, [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need(nBatches <= kIntMax, "batch too large");
        need_ptr(img, "img"); need_ptr(prj, "prj"); need_ptr(deg, "angles");
        std::vector<float> hdeg = host_copy(deg, nview);
        dd2d_fp_adjoint(img, prj, hdeg.data(), (int)nBatches, g, d, (int)nview, dsd, dso, type, fan, cur_stream());
        sync();
    });
}

CTP_EXPORT int ctpDistanceDrivenFanProjectionAdjoint(float* img, const float* prj, const float* deg, DD_GRID_ARGS,
                                                     float dsd, float dso, int typeProjector) {
    return dd_adjoint_entry("ctpDistanceDrivenFanProjectionAdjoint", img, prj, deg, nBatches, nx, ny, nz, dx, dy, dz,
                            cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso, typeProjector, true);
}
CTP_EXPORT int ctpDistanceDrivenParallelProjectionAdjoint(float* img, const float* prj, const float* deg, DD_GRID_ARGS,
                                                          int typeProjector) {
    return dd_adjoint_entry("ctpDistanceDrivenParallelProjectionAdjoint", img, prj, deg, nBatches, nx, ny, nz, dx, dy,
                            dz, cx, cy, cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, typeProjector, false);
}

// Persistent DD plans (no counterpart in the reference): the geometry tables of
// one fan / parallel geometry and batch count built once (angles in host
// memory); the calls are asynchronous on the given stream.
CTP_EXPORT int ctpDDPlanCreate(void** plan, int fan, const float* deg, DD_GRID_ARGS, float dsd, float dso,
                               int typeProjector) {
    return guarded("ctpDDPlanCreate", [&] {
        need_ptr(plan, "plan");
        *plan = nullptr;
        need(!empty(nBatches, nx, ny, nz, nu, nv, nview), "empty geometry");
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need(nBatches <= kIntMax && nBatches * nz <= kIntMax, "batch too large");
        need_ptr(deg, "angles");
        if (!fan) { dsd = 1000.f; dso = 500.f; }
        touch();
        *plan = dd_plan_create(deg, (int)nBatches, g, d, (int)nview, dsd, dso, typeProjector, fan != 0, cur_stream());
    });
}

CTP_EXPORT int ctpDDPlanForward(void* plan, float* prj, const float* img, void* stream) {
    return guarded("ctpDDPlanForward", [&] {
        need_ptr(plan, "plan"); need_ptr(prj, "prj"); need_ptr(img, "img");
        touch();
        dd_plan_forward(static_cast<DDPlan*>(plan), prj, img, static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpDDPlanBackward(void* plan, float* img, const float* prj, void* stream) {
    return guarded("ctpDDPlanBackward", [&] {
        need_ptr(plan, "plan"); need_ptr(img, "img"); need_ptr(prj, "prj");
        touch();
        dd_plan_backward(static_cast<DDPlan*>(plan), img, prj, static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpDDPlanAdjoint(void* plan, float* img, const float* prj, void* stream) {
    return guarded("ctpDDPlanAdjoint", [&] {
        need_ptr(plan, "plan"); need_ptr(img, "img"); need_ptr(prj, "prj");
        touch();
        dd_plan_adjoint(static_cast<DDPlan*>(plan), img, prj, static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpDDPlanDestroy(void* plan) {
    return guarded("ctpDDPlanDestroy", [&] {
        if (!plan) return;
        touch();
        dd_plan_destroy(static_cast<DDPlan*>(plan));
    });
}

// ===========================================================================
// Siddon fan beam (cuda/siddonFan.cu:128-274), host arrays
// ===========================================================================
template <bool FP>
static int siddon_fan_entry(const char* name, float* out, const float* in, const float* deg, int nBatchesArg,
                            size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                            size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v,
                            float dsd, float dso) {
    return guarded(name, [&] {
        need(nBatchesArg >= 0, "negative batch count");
        const size_t nBatches = (size_t)nBatchesArg;
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need_ptr(out, "out"); need_ptr(in, "in"); need_ptr(deg, "angles");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview;
        const size_t nin = FP ? nimg : nprj, nout = FP ? nprj : nimg;
        touch();
        DeviceBuffer di(nin * 4), dout(nout * 4), dd(nview * 4);
        up(di.p, in, nin * 4);
        up(dd.p, deg, nview * 4);
        zero(dout.p, nout * 4);
        if (FP) siddon_fan_fp(dout.as<float>(), di.as<float>(), dd.as<float>(), (int)nBatches, g, d, (int)nview, dsd, dso,
                              cur_stream());
        else siddon_fan_bp(dout.as<float>(), di.as<float>(), dd.as<float>(), (int)nBatches, g, d, (int)nview, dsd, dso,
                           cur_stream());
        sync();
        down(out, dout.p, nout * 4);
    });
}

// nBatches is an int here, as in cuda/siddonFan.cu:131,205 (the other DD 2D entry points take size_t)
#define SIDDON_FAN_ARGS                                                                                    \
    int nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, \
        size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v

CTP_EXPORT int cSiddonFanProjection(float* prj, const float* img, const float* deg, SIDDON_FAN_ARGS, float dsd,
                                    float dso) {
    return siddon_fan_entry<true>("cSiddonFanProjection", prj, img, deg, nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz,
                                  nu, nv, nview, da, dv, off_a, off_v, dsd, dso);
}
CTP_EXPORT int cSiddonFanBackprojection(float* img, const float* prj, const float* deg, SIDDON_FAN_ARGS, float dsd,
                                        float dso) {
    return siddon_fan_entry<false>("cSiddonFanBackprojection", img, prj, deg, nBatches, nx, ny, nz, dx, dy, dz, cx, cy,
                                   cz, nu, nv, nview, da, dv, off_a, off_v, dsd, dso);
}

// ===========================================================================
// Distance driven tomosynthesis (cuda/distanceDriven3D.cu:184-439)
// ===========================================================================
#define TOMO_ARGS                                                                                                 \
    const float *detCenter, const float *src, size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, \
        float dz, float cx, float cy, float cz, size_t nu, size_t nv, size_t nview, float du, float dv, float off_u, \
        float off_v, int typeProjector

template <bool FP, bool DEV>
static int tomo_entry(const char* name, float* out, const float* in, const float* detCenter, const float* src,
                      size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                      float cz, size_t nu, size_t nv, size_t nview, float du, float dv, float off_u, float off_v,
                      int type) {
    return guarded(name, [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, du, dv, off_u, off_v);
        need(nBatches <= kIntMax, "batch too large");
        need_ptr(out, "out"); need_ptr(in, "in"); need_ptr(detCenter, "detCenter"); need_ptr(src, "src");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview, ng = 3 * nview;
        if (DEV) {
            touch();
            if (FP) tomo_fp(out, in, detCenter, src, (int)nBatches, g, d, (int)nview, type, cur_stream());
            else tomo_bp(out, in, detCenter, src, (int)nBatches, g, d, (int)nview, type, cur_stream());
            sync();
        } else {
            const size_t nin = FP ? nimg : nprj, nout = FP ? nprj : nimg;
            touch();
            DeviceBuffer di(nin * 4), dout(nout * 4), dg(2 * ng * 4);
            float* geo = dg.as<float>();
            up(di.p, in, nin * 4);
            up(geo, detCenter, ng * 4);
            up(geo + ng, src, ng * 4);
            zero(dout.p, nout * 4);
            if (FP) tomo_fp(dout.as<float>(), di.as<float>(), geo, geo + ng, (int)nBatches, g, d, (int)nview, type, cur_stream());
            else tomo_bp(dout.as<float>(), di.as<float>(), geo, geo + ng, (int)nBatches, g, d, (int)nview, type, cur_stream());
            sync();
            down(out, dout.p, nout * 4);
        }
    });
}

CTP_EXPORT int cDistanceDrivenTomoProjection(float* prj, const float* img, TOMO_ARGS) {
    return tomo_entry<true, false>("cDistanceDrivenTomoProjection", prj, img, detCenter, src, nBatches, nx, ny, nz, dx,
                                   dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, typeProjector);
}
CTP_EXPORT int cupyDistanceDrivenTomoProjection(float* prj, const float* img, TOMO_ARGS) {
    return tomo_entry<true, true>("cupyDistanceDrivenTomoProjection", prj, img, detCenter, src, nBatches, nx, ny, nz, dx,
                                  dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, typeProjector);
}
CTP_EXPORT int cDistanceDrivenTomoBackprojection(float* img, const float* prj, TOMO_ARGS) {
    return tomo_entry<false, false>("cDistanceDrivenTomoBackprojection", img, prj, detCenter, src, nBatches, nx, ny, nz,
                                    dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, typeProjector);
}
CTP_EXPORT int cupyDistanceDrivenTomoBackprojection(float* img, const float* prj, TOMO_ARGS) {
    return tomo_entry<false, true>("cupyDistanceDrivenTomoBackprojection", img, prj, detCenter, src, nBatches, nx, ny,
                                   nz, dx, dy, dz, cx, cy, cz, nu, nv, nview, du, dv, off_u, off_v, typeProjector);
}

// ===========================================================================
// Ramp filters (cuda/fbp.cu:165-258, cuda/fbpParallel.cu:111-198)
// ===========================================================================
template <bool DEV>
static int filter_entry(const char* name, float* out, const float* in, int nBatches, size_t nu, size_t nv, size_t nview,
                        float da, float off_u, float dsd, float dso, int filter, bool equispace) {
    return guarded(name, [&] {
        need(nBatches >= 0, "negative batch");
        if (empty((size_t)nBatches, nu, nv, nview, 1, 1, 1)) return;
        need(nu <= 5000 && nu * nv <= kIntMax && nview <= kIntMax, "detector dimension too large");
        need(da != 0, "zero detector pitch");
        need(equispace || dso != 0, "zero dso");
        need_ptr(out, "out"); need_ptr(in, "in");
        const size_t n = (size_t)nBatches * nu * nv * nview;
        if (DEV) {
            touch();
            ramp_filter(out, in, nBatches, (int)nu, (int)nv, (int)nview, da, off_u, dsd, dso, filter, equispace, cur_stream());
            sync();
        } else {
            touch();
            DeviceBuffer di(n * 4), dout(n * 4);
            up(di.p, in, n * 4);
            ramp_filter(dout.as<float>(), di.as<float>(), nBatches, (int)nu, (int)nv, (int)nview, da, off_u, dsd, dso,
                        filter, equispace, cur_stream());
            sync();
            down(out, dout.p, n * 4);
        }
    });
}

CTP_EXPORT int cupyfbpFanFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview,
                                float da, float dv, float off_a, float off_v, float dsd, float dso, int typeFilter) {
    (void)dv; (void)off_v;
    return filter_entry<true>("cupyfbpFanFilter", pFPrj, pPrj, nBatches, nu, nv, nview, da, off_a, dsd, dso, typeFilter,
                              false);
}
CTP_EXPORT int cfbpFanFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview,
                             float da, float dv, float off_a, float off_v, float dsd, float dso, int typeFilter) {
    (void)dv; (void)off_v;
    return filter_entry<false>("cfbpFanFilter", pFPrj, pPrj, nBatches, nu, nv, nview, da, off_a, dsd, dso, typeFilter,
                               false);
}
CTP_EXPORT int cupyfbpParallelFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview,
                                     float du, float dv, float off_u, float off_v, int typeFilter) {
    (void)dv; (void)off_v;
    return filter_entry<true>("cupyfbpParallelFilter", pFPrj, pPrj, nBatches, nu, nv, nview, du, off_u, 0.f, 1.f,
                              typeFilter, true);
}
CTP_EXPORT int cfbpParallelFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview,
                                  float du, float dv, float off_u, float off_v, int typeFilter) {
    (void)dv; (void)off_v;
    return filter_entry<false>("cfbpParallelFilter", pFPrj, pPrj, nBatches, nu, nv, nview, du, off_u, 0.f, 1.f,
                               typeFilter, true);
}

// ===========================================================================
// Fan FBP backprojection (cuda/fbp.cu:362-481)
// ===========================================================================
template <bool DEV>
static int fbp_bp_entry(const char* name, float* img, const float* prj, const float* deg, size_t nBatches, size_t nx,
                        size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                        size_t nv, size_t nview, float da, float dv, float off_a, float off_v, float dsd, float dso,
                        bool fan = true) {
    return guarded(name, [&] {
        if (empty(nBatches, nx, ny, nz, nu, nv, nview)) return;
        Grid g = make_grid(nx, ny, nz, dx, dy, dz, cx, cy, cz);
        Det d = make_det(nu, nv, nview, da, dv, off_a, off_v);
        need(nBatches <= kIntMax, "batch too large");
        need_ptr(img, "img"); need_ptr(prj, "prj"); need_ptr(deg, "angles");
        const size_t nimg = nBatches * nx * ny * nz, nprj = nBatches * nu * nv * nview;
        if (DEV) {
            std::vector<float> hdeg = host_copy(deg, nview);
            fan_fbp_bp(img, prj, hdeg.data(), (int)nBatches, g, d, (int)nview, dsd, dso, cur_stream(), fan);
            sync();
        } else {
            touch();
            DeviceBuffer di(nimg * 4), dp(nprj * 4);
            up(dp.p, prj, nprj * 4);
            zero(di.p, nimg * 4);
            fan_fbp_bp(di.as<float>(), dp.as<float>(), deg, (int)nBatches, g, d, (int)nview, dsd, dso, cur_stream(), fan);
            sync();
            down(img, di.p, nimg * 4);
        }
    });
}

CTP_EXPORT int cupyfbpFanBackprojection(float* pImg, const float* pPrj, const float* pDeg, DD_GRID_ARGS, float dsd,
                                        float dso) {
    return fbp_bp_entry<true>("cupyfbpFanBackprojection", pImg, pPrj, pDeg, nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz,
                              nu, nv, nview, da, dv, off_a, off_v, dsd, dso);
}
// cuda/fbpTomo.cu:386  (device arrays; geometry rows (x, y, z))
CTP_EXPORT int cupyFbpTomoFilter(float* fprj, const float* prj, const float* detCenter, const float* src, int nBatches,
                                 size_t nu, size_t nv, size_t nview, float du, float dx, float dz, int typeFilter,
                                 float cutoffX, float cutoffZ) {
    return guarded("cupyFbpTomoFilter", [&] {
        if (nBatches <= 0 || nu == 0 || nv == 0 || nview == 0) return;
        need(nu <= kIntMax && nv <= kIntMax && nview <= 65535, "detector dimension too large");
        need((size_t)nBatches * nview * nv <= (size_t)0x7fffffff, "too many detector rows");
        need(du != 0, "zero detector pitch");
        need_ptr(fprj, "fprj"); need_ptr(prj, "prj"); need_ptr(detCenter, "detCenter"); need_ptr(src, "src");
        std::vector<float> hc = host_copy(detCenter, 3 * nview), hs = host_copy(src, 3 * nview);
        tomo_filter(fprj, prj, hc.data(), hs.data(), nBatches, (int)nu, (int)nv, (int)nview, du, dx, dz, typeFilter,
                    cutoffX, cutoffZ, cur_stream());
        sync();
    });
}

// cuda/fbpParallel.cu:273 (typeProjector is accepted and unused, as there)
CTP_EXPORT int cfbpParallelBackprojection(float* pImg, const float* pPrj, const float* pDeg, DD_GRID_ARGS,
                                          int typeProjector) {
    (void)typeProjector;
    return fbp_bp_entry<false>("cfbpParallelBackprojection", pImg, pPrj, pDeg, nBatches, nx, ny, nz, dx, dy, dz, cx, cy,
                               cz, nu, nv, nview, da, dv, off_a, off_v, 1000.f, 500.f, false);
}
CTP_EXPORT int cfbpFanBackprojection(float* pImg, const float* pPrj, const float* pDeg, DD_GRID_ARGS, float dsd,
                                     float dso) {
    return fbp_bp_entry<false>("cfbpFanBackprojection", pImg, pPrj, pDeg, nBatches, nx, ny, nz, dx, dy, dz, cx, cy, cz,
                               nu, nv, nview, da, dv, off_a, off_v, dsd, dso);
}

// ===========================================================================
// TV prior for SQS (the reference's libprior.so, prior-cuda/TotalVariation.cu)
// ===========================================================================
static void tv_dims(size_t B, size_t nx, size_t ny, size_t nz) {
    need(nx <= kIntMax && ny <= 65535 && nz <= 65535 && B <= 65535, "tv prior: dimension too large");
    need(nx * ny * nz <= kIntMax, "tv prior: more than 2^31-1 voxels per batch item");
}

template <bool DEV>
static int tv_entry(const char* name, float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz,
                    size_t B, size_t nx, size_t ny, size_t nz, float eps) {
    return guarded(name, [&] {
        if (B == 0 || nx == 0 || ny == 0 || nz == 0) return;
        tv_dims(B, nx, ny, nz);
        need_ptr(s1, "s1"); need_ptr(s2, "s2"); need_ptr(var, "var"); need_ptr(img, "img");
        const size_t n = B * nx * ny * nz;
        if (DEV) {
            touch();
            tv_sqs(s1, s2, var, img, wx, wy, wz, (int)B, (int)nx, (int)ny, (int)nz, eps, cur_stream());
            sync();
        } else {
            touch();
            DeviceBuffer a(n * 4), b(n * 4), c(n * 4), x(n * 4);
            up(x.p, img, n * 4);
            tv_sqs(a.as<float>(), b.as<float>(), c.as<float>(), x.as<float>(), wx, wy, wz, (int)B, (int)nx, (int)ny,
                   (int)nz, eps, cur_stream());
            sync();
            down(s1, a.p, n * 4);
            down(s2, b.p, n * 4);
            down(var, c.p, n * 4);
        }
    });
}

CTP_EXPORT int cupyTVSQS3D(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz,
                           size_t nBatches, size_t nx, size_t ny, size_t nz, float eps) {
    return tv_entry<true>("cupyTVSQS3D", s1, s2, var, img, wx, wy, wz, nBatches, nx, ny, nz, eps);
}
CTP_EXPORT int cTVSQS3D(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz,
                        size_t nBatches, size_t nx, size_t ny, size_t nz, float eps) {
    return tv_entry<false>("cTVSQS3D", s1, s2, var, img, wx, wy, wz, nBatches, nx, ny, nz, eps);
}

// One SQS step with the TV prior, fused (device pointers):
//   out = img - (acc * bp + beta * s1(img)) / (norm + beta * s2(img)); var(img) if var != NULL
CTP_EXPORT int ctpTVSQSUpdate3D(float* out, float* var, const float* img, const float* bp, const float* norm, float beta,
                                float acc, float wx, float wy, float wz, size_t nBatches, size_t nx, size_t ny,
                                size_t nz, float eps) {
    return guarded("ctpTVSQSUpdate3D", [&] {
        if (nBatches == 0 || nx == 0 || ny == 0 || nz == 0) return;
        tv_dims(nBatches, nx, ny, nz);
        need_ptr(out, "out"); need_ptr(img, "img"); need_ptr(bp, "bp"); need_ptr(norm, "norm");
        touch();
        tv_sqs_update(out, var, img, bp, norm, beta, acc, wx, wy, wz, (int)nBatches, (int)nx, (int)ny, (int)nz, eps,
                      cur_stream());
        sync();
    });
}

CTP_EXPORT int ctpTVSQSBackward3D(float* dimg, const float* img, const float* g1, const float* g2, const float* gvar,
                                  float wx, float wy, float wz, size_t nBatches, size_t nx, size_t ny, size_t nz,
                                  float eps) {
    return guarded("ctpTVSQSBackward3D", [&] {
        if (nBatches == 0 || nx == 0 || ny == 0 || nz == 0) return;
        tv_dims(nBatches, nx, ny, nz);
        need_ptr(dimg, "dimg"); need_ptr(img, "img"); need_ptr(g1, "g1"); need_ptr(g2, "g2");
        need(dimg != img && dimg != g1 && dimg != g2 && dimg != gvar, "ctpTVSQSBackward3D: dimg must not alias an input");
        touch();
        tv_sqs_backward(dimg, img, g1, g2, gvar, wx, wy, wz, (int)nBatches, (int)nx, (int)ny, (int)nz, eps,
                        cur_stream());
        sync();
    });
}

// Asynchronous forms of cupyTVSQS3D and ctpTVSQSBackward3D (no counterpart in
// the reference): the kernels are queued on 'stream' and the call returns
// without waiting, so a caller's loop (the differentiable SQS-TV network)
// keeps the device busy while the host queues the next work.
CTP_EXPORT int ctpTVSQS3DAsync(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz,
                               size_t nBatches, size_t nx, size_t ny, size_t nz, float eps, void* stream) {
    return guarded("ctpTVSQS3DAsync", [&] {
        if (nBatches == 0 || nx == 0 || ny == 0 || nz == 0) return;
        tv_dims(nBatches, nx, ny, nz);
        need_ptr(s1, "s1"); need_ptr(s2, "s2"); need_ptr(var, "var"); need_ptr(img, "img");
        touch();
        tv_sqs(s1, s2, var, img, wx, wy, wz, (int)nBatches, (int)nx, (int)ny, (int)nz, eps,
               static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpTVSQSBackward3DAsync(float* dimg, const float* img, const float* g1, const float* g2,
                                       const float* gvar, float* scratch, float wx, float wy, float wz,
                                       size_t nBatches, size_t nx, size_t ny, size_t nz, float eps, void* stream) {
    return guarded("ctpTVSQSBackward3DAsync", [&] {
        if (nBatches == 0 || nx == 0 || ny == 0 || nz == 0) return;
        tv_dims(nBatches, nx, ny, nz);
        need_ptr(dimg, "dimg"); need_ptr(img, "img"); need_ptr(g1, "g1"); need_ptr(g2, "g2");
        need(dimg != img && dimg != g1 && dimg != g2 && dimg != gvar,
             "ctpTVSQSBackward3DAsync: dimg must not alias an input");
        need(scratch == nullptr || (reinterpret_cast<size_t>(scratch) % 8 == 0 && scratch != dimg),
             "ctpTVSQSBackward3DAsync: scratch must be 8-byte aligned and distinct from dimg");
        touch();
        tv_sqs_backward(dimg, img, g1, g2, gvar, wx, wy, wz, (int)nBatches, (int)nx, (int)ny, (int)nz, eps,
                        static_cast<hipStream_t>(stream), scratch);
    });
}

// ===========================================================================
// Batched cone-beam plans (no counterpart in the reference)
// ===========================================================================
struct ctpConeJob {
    size_t nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
    size_t nu, nv, nview;
    float du, dv, off_u, off_v;
    const float* det_center;
    const float* det_u;
    const float* det_v;
    const float* src;
    float* img;
    float* prj;
};

CTP_EXPORT int ctpConePlanCreate(void** plan, const ctpConeJob* jobs, int njobs, int autotune) {
    return guarded("ctpConePlanCreate", [&] {
        need_ptr(plan, "plan");
        *plan = nullptr;
        need(njobs > 0 && jobs != nullptr, "no jobs");
        std::vector<ConeJobSpec> specs;
        for (int j = 0; j < njobs; j++) {
            const ctpConeJob& J = jobs[j];
            need(!empty(1, J.nx, J.ny, J.nz, J.nu, J.nv, J.nview), "empty job");
            ConeJobSpec S{};
            S.g = make_grid(J.nx, J.ny, J.nz, J.dx, J.dy, J.dz, J.cx, J.cy, J.cz);
            S.d = make_det(J.nu, J.nv, J.nview, J.du, J.dv, J.off_u, J.off_v);
            S.nview = (int)J.nview;
            need_ptr(J.det_center, "det_center"); need_ptr(J.det_u, "det_u");
            need_ptr(J.det_v, "det_v"); need_ptr(J.src, "src");
            need_ptr(J.img, "img"); need_ptr(J.prj, "prj");
            S.det_center = J.det_center; S.det_u = J.det_u; S.det_v = J.det_v; S.src = J.src;
            S.img = J.img;
            S.prj = J.prj;
            specs.push_back(S);
        }
        touch();
        *plan = new ConePlan(specs, autotune != 0);
    });
}

CTP_EXPORT int ctpConePlanForward(void* plan, void* stream) {
    return guarded("ctpConePlanForward", [&] {
        need_ptr(plan, "plan");
        touch();
        static_cast<ConePlan*>(plan)->forward(static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpConePlanBackward(void* plan, void* stream) {
    return guarded("ctpConePlanBackward", [&] {
        need_ptr(plan, "plan");
        touch();
        static_cast<ConePlan*>(plan)->backward(static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpConePlanConfig(void* plan, int* tile_u, int* xcd_order, int* zrun) {
    return guarded("ctpConePlanConfig", [&] {
        need_ptr(plan, "plan");
        int a, b, c;
        static_cast<ConePlan*>(plan)->config(a, b, c);
        if (tile_u) *tile_u = a;
        if (xcd_order) *xcd_order = b;
        if (zrun) *zrun = c;
    });
}

CTP_EXPORT int ctpConePlanDestroy(void* plan) {
    return guarded("ctpConePlanDestroy", [&] {
        if (!plan) return;
        touch();
        CTP_CHECK(hipDeviceSynchronize());
        delete static_cast<ConePlan*>(plan);
    });
}

// ===========================================================================
// Batched plans of uneven distance-driven jobs (no counterpart in the reference)
// ===========================================================================
struct ctpDDJob {
    int fan;
    int typeProjector;
    size_t nBatches, nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
    size_t nu, nv, nview;
    float da, dv, off_a, off_v;
    float dsd, dso;
    const float* deg;
    float* img;
    float* prj;
};

CTP_EXPORT int ctpDDBatchCreate(void** batch, const ctpDDJob* jobs, int njobs, int autotune) {
    return guarded("ctpDDBatchCreate", [&] {
        need_ptr(batch, "batch");
        *batch = nullptr;
        need(njobs > 0 && jobs != nullptr, "no jobs");
        std::vector<DDBatchJobSpec> specs;
        for (int j = 0; j < njobs; j++) {
            const ctpDDJob& J = jobs[j];
            need(!empty(J.nBatches, J.nx, J.ny, J.nz, J.nu, J.nv, J.nview), "empty job");
            need(J.nBatches <= kIntMax && J.nBatches * J.nz <= kIntMax, "batch too large");
            need_ptr(J.deg, "angles");
            need_ptr(J.img, "img");
            need_ptr(J.prj, "prj");
            DDBatchJobSpec S{};
            S.fan = J.fan != 0;
            S.type = J.typeProjector;
            S.B = (int)J.nBatches;
            S.g = make_grid(J.nx, J.ny, J.nz, J.dx, J.dy, J.dz, J.cx, J.cy, J.cz);
            S.d = make_det(J.nu, J.nv, J.nview, J.da, J.dv, J.off_a, J.off_v);
            S.nview = (int)J.nview;
            // the parallel-beam kernels take no source distances (the single
            // parallel entry points pass these placeholders too)
            S.dsd = S.fan ? J.dsd : 1000.f;
            S.dso = S.fan ? J.dso : 500.f;
            S.deg = J.deg;
            S.img = J.img;
            S.prj = J.prj;
            specs.push_back(S);
        }
        touch();
        *batch = new DDBatchPlan(specs, autotune != 0);
    });
}

CTP_EXPORT int ctpDDBatchForward(void* batch, void* stream) {
    return guarded("ctpDDBatchForward", [&] {
        need_ptr(batch, "batch");
        touch();
        static_cast<DDBatchPlan*>(batch)->forward(static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpDDBatchBackward(void* batch, void* stream) {
    return guarded("ctpDDBatchBackward", [&] {
        need_ptr(batch, "batch");
        touch();
        static_cast<DDBatchPlan*>(batch)->backward(static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpDDBatchConfig(void* batch, int* fp_lanes, int* bp_lanes, int* fp_persistent, int* bp_persistent,
                                int* fp_streams, int* bp_streams, int* fp_launches, int* bp_launches) {
    return guarded("ctpDDBatchConfig", [&] {
        need_ptr(batch, "batch");
        const DDBatchPlan* b = static_cast<const DDBatchPlan*>(batch);
        int a = 0, c = 0, e = 0, f = 0, g = 0, h = 0;
        b->config(a, c, e, f, g, h);
        if (fp_lanes) *fp_lanes = a;
        if (bp_lanes) *bp_lanes = c;
        if (fp_persistent) *fp_persistent = e;
        if (bp_persistent) *bp_persistent = f;
        if (fp_streams) *fp_streams = g;
        if (bp_streams) *bp_streams = h;
        if (fp_launches) *fp_launches = b->launches(true);
        if (bp_launches) *bp_launches = b->launches(false);
    });
}

// ctpFilterJob of include/ctprojector.h
struct ctpFilterJob {
    int fan;
    int typeFilter;
    size_t nBatches, nu, nv, nview;
    float da, off_a, dsd, dso;
    const float* prj;
    float* fprj;
};

CTP_EXPORT int ctpFilterBatchCreate(void** batch, const ctpFilterJob* jobs, int njobs) {
    return guarded("ctpFilterBatchCreate", [&] {
        need_ptr(batch, "batch");
        *batch = nullptr;
        need(njobs > 0 && jobs != nullptr, "no jobs");
        std::vector<FilterBatchJobSpec> specs;
        for (int j = 0; j < njobs; j++) {
            const ctpFilterJob& J = jobs[j];
            need(!empty(J.nBatches, J.nu, J.nv, J.nview, 1, 1, 1), "empty job");
            need(J.nu <= 5000, "filter batch: detector rows of more than 5000 columns (as the single-call filter)");
            need(J.nBatches <= kIntMax && J.nv <= kIntMax && J.nview <= kIntMax, "job too large");
            need(J.da != 0, "zero detector pitch");
            need(!J.fan || J.dso != 0, "zero dso");
            need_ptr(J.prj, "prj");
            need_ptr(J.fprj, "fprj");
            FilterBatchJobSpec S{};
            S.B = (int)J.nBatches;
            S.nu = (int)J.nu;
            S.nv = (int)J.nv;
            S.nview = (int)J.nview;
            S.da = J.da;
            S.off_u = J.off_a;
            S.dso = J.fan ? J.dso : 1.f;
            S.filter = J.typeFilter;
            S.equispace = J.fan == 0;
            S.in = J.prj;
            S.out = J.fprj;
            specs.push_back(S);
        }
        touch();
        *batch = new FilterBatchPlan(specs);
    });
}

CTP_EXPORT int ctpFilterBatchRun(void* batch, void* stream) {
    return guarded("ctpFilterBatchRun", [&] {
        need_ptr(batch, "batch");
        touch();
        static_cast<FilterBatchPlan*>(batch)->run(static_cast<hipStream_t>(stream));
    });
}

CTP_EXPORT int ctpFilterBatchDestroy(void* batch) {
    return guarded("ctpFilterBatchDestroy", [&] {
        if (!batch) return;
        touch();
        CTP_CHECK(hipDeviceSynchronize());
        delete static_cast<FilterBatchPlan*>(batch);
    });
}

CTP_EXPORT int ctpDDBatchDestroy(void* batch) {
    return guarded("ctpDDBatchDestroy", [&] {
        if (!batch) return;
        touch();
        CTP_CHECK(hipDeviceSynchronize());
        delete static_cast<DDBatchPlan*>(batch);
    });
}
