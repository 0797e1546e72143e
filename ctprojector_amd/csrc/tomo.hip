// Distance-driven tomosynthesis projector (z main axis, detector axes fixed to
// u = (1,0,0), v = (0,1,0)) on MI355X (gfx950): cuda/distanceDriven3D.cu,
// cuda/distanceDriven3DBoxInt.cu, cuda/distanceDriven3DBranchless.cu,
// cuda/distanceDriven3DUtils.cu of the reference.
//
// FP: one lane per detector cell; for every image plane z = iz (voxel edge in
// index space) the four edge-midpoint rays of the cell span a box whose area
// average is accumulated (IntegralBoxXY, cuda/cudaMath.cu:336-397), then scaled
// by the ray length through one slab.
// BP: one lane per voxel; for every view the voxel's four edge midpoints are
// projected onto the detector (centre-indexed, cuda/distanceDriven3DUtils.cu:45-56)
// and the area average of the pre-weighted projections over that box is
// accumulated.
//
// The reference's branchless mode evaluates the same box integrals through
// FP64 summed-area tables (a summed-area table of a piecewise-constant image is
// exactly bilinear between its nodes); here both modes evaluate the box
// integral directly.  Batch items are handled independently (the reference's
// BoxInt BP reuses batch 0 and its branchless BP pre-weights view 0 only,
// cuda/distanceDriven3DBoxInt.cu:257-259, cuda/distanceDriven3DBranchless.cu:347).
#include "common.h"
#include "runtime.h"
#include "ops.h"

#include <cmath>

namespace ctp {

// exact integral of a piecewise-constant plane over [x1,x2]x[y1,y2] clamped to
// the plane (returns 0 for an empty box)
__device__ inline float box_xy(const float* __restrict__ p, float x1, float y1, float x2, float y2, int nx, int ny) {
    const float a = clampf(x1, 0.f, (float)nx), b = clampf(x2, 0.f, (float)nx);
    const float c = clampf(y1, 0.f, (float)ny), e = clampf(y2, 0.f, (float)ny);
    if (!(b > a) || !(e > c)) return 0.f;
    const int i1 = min((int)a, nx - 1), i2 = min((int)b, nx - 1);
    const int j1 = min((int)c, ny - 1), j2 = min((int)e, ny - 1);
    float s = 0.f;
    for (int j = j1; j <= j2; j++) {
        const float wy = fminf(e, (float)(j + 1)) - fmaxf(c, (float)j);
        float r = 0.f;
        for (int i = i1; i <= i2; i++) r = fmaf(fminf(b, (float)(i + 1)) - fmaxf(a, (float)i), p[(size_t)j * nx + i], r);
        s = fmaf(wy, r, s);
    }
    return s;
}

__device__ inline float3 ld3t(const float* p, int i) { return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

__global__ __launch_bounds__(256) void tomo_fp_kernel(float* __restrict__ prj, const float* __restrict__ img,
                                                      const float* __restrict__ pdc, const float* __restrict__ psrc,
                                                      Grid g, Det d, int nview) {
    const int iu = blockIdx.x * 32 + (threadIdx.x & 31);
    const int iv = blockIdx.y * 8 + (threadIdx.x >> 5);
    const int view = blockIdx.z;
    if (iu >= d.nu || iv >= d.nv) return;
    // cuda/distanceDriven3DBoxInt.cu:56-113
    const float u = ((float)iu - d.off_u - (d.nu - 1) / 2.0f) * d.du;
    const float v = ((float)iv - d.off_v - (d.nv - 1) / 2.0f) * d.dv;
    const float3 C = ld3t(pdc, view);
    const float3 S = ld3t(psrc, view);
    auto toimg = [&](float3 p) {
        return make_float3((p.x - g.cx) / g.dx + g.nx / 2.0f, (p.y - g.cy) / g.dy + g.ny / 2.0f,
                           (p.z - g.cz) / g.dz + g.nz / 2.0f);
    };
    float3 x1 = toimg(make_float3(C.x + u - d.du / 2, C.y + v, C.z));
    float3 x2 = toimg(make_float3(C.x + u + d.du / 2, C.y + v, C.z));
    float3 y1 = toimg(make_float3(C.x + u, C.y + v - d.dv / 2, C.z));
    float3 y2 = toimg(make_float3(C.x + u, C.y + v + d.dv / 2, C.z));
    const float3 s = toimg(S);
    if (x1.x > x2.x) { const float3 t = x1; x1 = x2; x2 = t; }
    if (y1.y > y2.y) { const float3 t = y1; y1 = y2; y2 = t; }
    const float rx1 = (x1.x - s.x) / (x1.z - s.z), rx2 = (x2.x - s.x) / (x2.z - s.z);
    const float ry1 = (y1.y - s.y) / (y1.z - s.z), ry2 = (y2.y - s.y) / (y2.z - s.z);
    float val = 0.f;
    const size_t npl = (size_t)g.nx * g.ny;
    for (int iz = 0; iz < g.nz; iz++) {
        const float h = (float)iz - s.z;
        const float a1 = s.x + rx1 * h, a2 = s.x + rx2 * h;
        const float b1 = s.y + ry1 * h, b2 = s.y + ry2 * h;
        val += box_xy(img + iz * npl, a1, b1, a2, b2, g.nx, g.ny) / ((a2 - a1) * (b2 - b1));
    }
    const float3 D = make_float3(C.x + u, C.y + v, C.z);
    val *= g.dz / fabsf(S.z - D.z) *
           sqrtf((S.z - D.z) * (S.z - D.z) + (S.y - D.y) * (S.y - D.y) + (S.x - D.x) * (S.x - D.x));
    prj[((size_t)view * d.nv + iv) * d.nu + iu] = val;
}

void tomo_fp(float* prj, const float* img, const float* det_center, const float* src, int B, const Grid& g,
             const Det& d, int nview, int type, hipStream_t stream) {
    (void)type;
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    dim3 grid(ceil_div(d.nu, 32), ceil_div(d.nv, 8), nview), block(256);
    for (int b = 0; b < B; b++) {
        TimedLaunch t("tomo_fp", stream);
        tomo_fp_kernel<<<grid, block, 0, stream>>>(prj + b * nprj, img + b * nimg, det_center, src, g, d, nview);
        CTP_LAUNCH_CHECK();
    }
}

// PreweightBPCartKernelXY, cuda/distanceDriven3DUtils.cu:10-40
__global__ __launch_bounds__(256) void tomo_preweight_kernel(float* __restrict__ w, const float* __restrict__ prj,
                                                             const float* __restrict__ pdc, const float* __restrict__ psrc,
                                                             float dz, Det d, int nview) {
    const long n = (long)d.nu * d.nv * nview;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int iu = (int)(i % d.nu);
    const int iv = (int)((i / d.nu) % d.nv);
    const int view = (int)(i / ((long)d.nu * d.nv));
    const float u = ((float)iu - d.off_u - (d.nu - 1) / 2.0f) * d.du;
    const float v = ((float)iv - d.off_v - (d.nv - 1) / 2.0f) * d.dv;
    const float3 C = ld3t(pdc, view), S = ld3t(psrc, view);
    const float3 D = make_float3(C.x + u, C.y + v, C.z);
    w[i] = prj[i] * (dz / fabsf(S.z - D.z) *
                     sqrtf((S.z - D.z) * (S.z - D.z) + (S.y - D.y) * (S.y - D.y) + (S.x - D.x) * (S.x - D.x)));
}

__global__ __launch_bounds__(256) void tomo_bp_kernel(float* __restrict__ img, const float* __restrict__ w,
                                                      const float* __restrict__ pdc, const float* __restrict__ psrc,
                                                      Grid g, Det d, int nview) {
    const int ix = blockIdx.x * 32 + (threadIdx.x & 31);
    const int iy = blockIdx.y * 8 + (threadIdx.x >> 5);
    const int iz = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    // voxel centre, cuda/distanceDriven3DBoxInt.cu:191-193
    const float x = ((float)ix - (g.nx - 1) / 2.0f) * g.dx + g.cx;
    const float y = ((float)iy - (g.ny - 1) / 2.0f) * g.dy + g.cy;
    const float z = ((float)iz - (g.nz - 1) / 2.0f) * g.dz + g.cz;
    const float ou = d.off_u + (d.nu - 1) / 2.0f, ov = d.off_v + (d.nv - 1) / 2.0f;
    const size_t nuv = (size_t)d.nu * d.nv;
    float val = 0.f;
    for (int view = 0; view < nview; view++) {
        const float3 S = ld3t(psrc, view), C = ld3t(pdc, view);
        // ProjectConeToDetCart, cuda/distanceDriven3DUtils.cu:45-56
        const float r = (C.z - S.z) / (z - S.z);
        const float u1 = (S.x - C.x + r * (x - g.dx / 2 - S.x)) / d.du + ou;
        const float u2 = (S.x - C.x + r * (x + g.dx / 2 - S.x)) / d.du + ou;
        const float v1 = (S.y - C.y + r * (y - g.dy / 2 - S.y)) / d.dv + ov;
        const float v2 = (S.y - C.y + r * (y + g.dy / 2 - S.y)) / d.dv + ov;
        val += box_xy(w + view * nuv, u1, v1, u2, v2, d.nu, d.nv) / ((u2 - u1) * (v2 - v1));
    }
    img[((size_t)iz * g.ny + iy) * g.nx + ix] = val;
}

void tomo_bp(float* img, const float* prj, const float* det_center, const float* src, int B, const Grid& g,
             const Det& d, int nview, int type, hipStream_t stream) {
    (void)type;
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>(nprj);
    ws.reserve(pl.bytes);
    float* w = ws.take_n<float>(nprj);
    dim3 grid(ceil_div(g.nx, 32), ceil_div(g.ny, 8), g.nz), block(256);
    for (int b = 0; b < B; b++) {
        tomo_preweight_kernel<<<ceil_div((long)nprj, 256), 256, 0, stream>>>(w, prj + b * nprj, det_center, src, g.dz, d, nview);
        CTP_LAUNCH_CHECK();
        TimedLaunch t("tomo_bp", stream);
        tomo_bp_kernel<<<grid, block, 0, stream>>>(img + b * nimg, w, det_center, src, g, d, nview);
        CTP_LAUNCH_CHECK();
    }
}

}  // namespace ctp
