// Anisotropic TV prior for SQS on MI355X (gfx950): the reference's
// prior-cuda/TotalVariation.cu (Variation3DKernel :21-55, TVSQS3DKernel
// :57-139, TVSQS3D_gpu :141-180).
//
// The reference runs two kernels per batch item: one writes the variation map
// var, the next reads img and var again to form s1 and s2 (24 B of HBM traffic
// per voxel).  Here one pass computes var at the voxel and at its three
// forward neighbours from the image and writes s1, s2, var once (16 B per
// voxel).  tv_sqs_update fuses the SQS step
// x <- x - (a * bp + beta * s1) / (norm + beta * s2) into the same pass, so s1
// and s2 never reach memory (img, bp, norm in, x and var out).  The default
// kernels march along z through LDS-staged slices (tv_zmarch_kernel); the
// one-lane-per-voxel kernels (neighbours from cache) remain as tv_mode 2.
#include "common.h"
#include "runtime.h"
#include "ops.h"

namespace ctp {

struct TVW {
    float wx, wy, wz, eps;
};

// variation at voxel (ix, iy, iz) of one volume: the reference's
// Variation3DKernel with zero differences at the lower edges
__device__ inline float tv_var(const float* __restrict__ img, size_t ind, int ix, int iy, int iz, int nx, size_t nxy,
                               const TVW& w) {
    const float x = img[ind];
    const float dx = ix > 0 ? x - img[ind - 1] : 0.f;
    const float dy = iy > 0 ? x - img[ind - nx] : 0.f;
    const float dz = iz > 0 ? x - img[ind - nxy] : 0.f;
    return sqrtf(w.wx * dx * dx + w.wy * dy * dy + w.wz * dz * dz + w.eps);
}

// s1, s2 of the reference's TVSQS3DKernel (the value at the upper edges uses
// the voxel's own variation for the missing neighbour, as there)
__device__ inline void tv_s12(const float* __restrict__ img, size_t ind, int ix, int iy, int iz, int nx, int ny,
                              int nz, size_t nxy, const TVW& w, float& s1, float& s2, float& v) {
    const float x = img[ind];
    v = tv_var(img, ind, ix, iy, iz, nx, nxy, w);
    float dx0 = ix > 0 ? x - img[ind - 1] : 0.f;
    float dy0 = iy > 0 ? x - img[ind - nx] : 0.f;
    float dz0 = iz > 0 ? x - img[ind - nxy] : 0.f;
    float dx1 = 0.f, dy1 = 0.f, dz1 = 0.f, varx = v, vary = v, varz = v;
    if (ix < nx - 1) { dx1 = x - img[ind + 1]; varx = tv_var(img, ind + 1, ix + 1, iy, iz, nx, nxy, w); }
    if (iy < ny - 1) { dy1 = x - img[ind + nx]; vary = tv_var(img, ind + nx, ix, iy + 1, iz, nx, nxy, w); }
    if (iz < nz - 1) { dz1 = x - img[ind + nxy]; varz = tv_var(img, ind + nxy, ix, iy, iz + 1, nx, nxy, w); }
    dx0 *= w.wx; dx1 *= w.wx;
    dy0 *= w.wy; dy1 *= w.wy;
    dz0 *= w.wz; dz1 *= w.wz;
    s1 = (dx0 + dy0 + dz0) / v + dx1 / varx + dy1 / vary + dz1 / varz;
    s2 = (w.wx + w.wy + w.wz) / v + w.wx / varx + w.wy / vary + w.wz / varz;
}

// grid: (ceil(nx/64), ny, nz * B); one lane per voxel along x
__global__ __launch_bounds__(64) void tv_sqs_kernel(float* __restrict__ s1, float* __restrict__ s2,
                                                    float* __restrict__ var, const float* __restrict__ img, int nx,
                                                    int ny, int nz, TVW w) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y;
    const int b = blockIdx.z / nz, iz = blockIdx.z - b * nz;
    if (ix >= nx) return;
    const size_t nxy = (size_t)nx * ny;
    const size_t base = (size_t)b * nxy * nz;
    const size_t ind = (size_t)iz * nxy + (size_t)iy * nx + ix;
    float a, c, v;
    tv_s12(img + base, ind, ix, iy, iz, nx, ny, nz, nxy, w, a, c, v);
    s1[base + ind] = a;
    s2[base + ind] = c;
    var[base + ind] = v;
}

// out = img - (acc * bp + beta * s1) / (norm + beta * s2); var optional
__global__ __launch_bounds__(64) void tv_sqs_update_kernel(float* __restrict__ out, float* __restrict__ var,
                                                           const float* __restrict__ img, const float* __restrict__ bp,
                                                           const float* __restrict__ norm, float beta, float acc,
                                                           int nx, int ny, int nz, TVW w) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y;
    const int b = blockIdx.z / nz, iz = blockIdx.z - b * nz;
    if (ix >= nx) return;
    const size_t nxy = (size_t)nx * ny;
    const size_t base = (size_t)b * nxy * nz;
    const size_t ind = (size_t)iz * nxy + (size_t)iy * nx + ix;
    float a, c, v;
    tv_s12(img + base, ind, ix, iy, iz, nx, ny, nz, nxy, w, a, c, v);
    const size_t i = base + ind;
    out[i] = img[i] - (bp[i] * acc + beta * a) / (norm[i] + beta * c);
    if (var) var[i] = v;
}

// ---------------------------------------------------------------------------
// gradient of the surrogate (no counterpart in the reference): for the
// unrolled, differentiable SQS-TV network, dimg = (ds1/dimg)^T g1 +
// (ds2/dimg)^T g2 + (dvar/dimg)^T gv, as two deterministic gathers.
//
// With D0_d(q) = x(q) - x(q - e_d) (0 at the lower edge), A(q) = sum_d w_d D0_d(q),
// W = wx + wy + wz and v the variation, the loss gradient with respect to v(q) is
//   Gv(q) = gv(q) - [g1(q) A(q) + g2(q) W
//                    + sum_{d: q_d > 0} w_d (g2(q - e_d) - g1(q - e_d) D0_d(q))
//                    + sum_{d: q_d = n_d - 1} w_d g2(q)] / v(q)^2
// (the last sum: at the upper edge s1, s2 use v(q) for the missing neighbour).
// Pass 1 stores P = Gv / v and R = 1 / v; pass 2 gathers
//   dimg(q) = sum_{d: q_d > 0} w_d (D0_d(q) P(q) + (g1(q) - g1(q - e_d)) R(q))
//           + sum_{d: q_d < n_d - 1} w_d ((g1(q) - g1(q + e_d)) R(q + e_d) - D0_d(q + e_d) P(q + e_d)).
__global__ __launch_bounds__(64) void tv_grad_pr_kernel(float2* __restrict__ pr, const float* __restrict__ img,
                                                        const float* __restrict__ g1, const float* __restrict__ g2,
                                                        const float* __restrict__ gv, int nx, int ny, int nz, TVW w) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y;
    const int b = blockIdx.z / nz, iz = blockIdx.z - b * nz;
    if (ix >= nx) return;
    const size_t nxy = (size_t)nx * ny;
    const size_t base = (size_t)b * nxy * nz;
    const size_t q = base + (size_t)iz * nxy + (size_t)iy * nx + ix;
    const float x = img[q];
    const float dx = ix > 0 ? x - img[q - 1] : 0.f;
    const float dy = iy > 0 ? x - img[q - nx] : 0.f;
    const float dz = iz > 0 ? x - img[q - nxy] : 0.f;
    const float v = sqrtf(w.wx * dx * dx + w.wy * dy * dy + w.wz * dz * dz + w.eps);
    const float a1 = g1[q], a2 = g2[q];
    float num = a1 * (w.wx * dx + w.wy * dy + w.wz * dz) + a2 * (w.wx + w.wy + w.wz);
    if (ix > 0) num += w.wx * (g2[q - 1] - g1[q - 1] * dx);
    if (iy > 0) num += w.wy * (g2[q - nx] - g1[q - nx] * dy);
    if (iz > 0) num += w.wz * (g2[q - nxy] - g1[q - nxy] * dz);
    if (ix == nx - 1) num += w.wx * a2;
    if (iy == ny - 1) num += w.wy * a2;
    if (iz == nz - 1) num += w.wz * a2;
    const float r = 1.f / v;
    const float gvq = (gv ? gv[q] : 0.f) - num * r * r;
    pr[q] = make_float2(gvq * r, r);
}

__global__ __launch_bounds__(64) void tv_grad_x_kernel(float* __restrict__ dimg, const float2* __restrict__ pr,
                                                       const float* __restrict__ img, const float* __restrict__ g1,
                                                       int nx, int ny, int nz, TVW w) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y;
    const int b = blockIdx.z / nz, iz = blockIdx.z - b * nz;
    if (ix >= nx) return;
    const size_t nxy = (size_t)nx * ny;
    const size_t base = (size_t)b * nxy * nz;
    const size_t q = base + (size_t)iz * nxy + (size_t)iy * nx + ix;
    const float x = img[q], a = g1[q];
    const float2 own = pr[q];
    float sum = 0.f;
    // one axis: step s in the flattened index, weight wd, lower / upper neighbour present
    auto axis = [&](size_t s, float wd, bool lower, bool upper) {
        if (lower) sum += wd * ((x - img[q - s]) * own.x + (a - g1[q - s]) * own.y);
        if (upper) {
            const float2 nb = pr[q + s];
            sum += wd * ((a - g1[q + s]) * nb.y - (img[q + s] - x) * nb.x);
        }
    };
    axis(1, w.wx, ix > 0, ix < nx - 1);
    axis((size_t)nx, w.wy, iy > 0, iy < ny - 1);
    axis(nxy, w.wz, iz > 0, iz < nz - 1);
    dimg[q] = sum;
}

// ---------------------------------------------------------------------------
// z-marching form of tv_sqs / tv_sqs_update: a block owns a 64 x 4 column tile
// of one volume and walks up to 32 slices of it.  Each image slice is read
// from memory once per tile (with a one-voxel halo) into LDS, the variation of
// the next slice is computed once per voxel (tile plus its upper halo) into
// LDS, and the voxel's own values along z stay in registers: per voxel one
// 4-byte read (plus halo) and the writes, instead of the per-voxel kernels'
// thirteen neighbour reads.  Arithmetic per voxel is that of tv_s12.
constexpr int kTVX = 64, kTVY = 4, kTVZ = 32;

template <bool UPDATE>
__global__ __launch_bounds__(256) void tv_zmarch_kernel(float* __restrict__ o1, float* __restrict__ o2,
                                                        float* __restrict__ ovar, const float* __restrict__ img,
                                                        const float* __restrict__ bp, const float* __restrict__ norm,
                                                        float beta, float acc, int nx, int ny, int nz, int nzc,
                                                        TVW w) {
    __shared__ float im[2][kTVY + 2][kTVX + 2];   // image slices z, z+1 (ring), halo -1..+1
    __shared__ float va[2][kTVY + 1][kTVX + 1];   // variation slices z, z+1 (ring), tile + upper halo
    const int tx = threadIdx.x & (kTVX - 1), ty = threadIdx.x / kTVX;
    const int x0 = blockIdx.x * kTVX, y0 = blockIdx.y * kTVY;
    const int b = blockIdx.z / nzc, zb = (blockIdx.z - b * nzc) * kTVZ;
    const int ze = min(nz, zb + kTVZ);
    const size_t nxy = (size_t)nx * ny;
    const size_t base = (size_t)b * nxy * nz;
    const float* __restrict__ vol = img + base;
    const int ix = x0 + tx, iy = y0 + ty;
    const bool own = ix < nx && iy < ny;

    // a slice with its halo is (kTVY + 2) * (kTVX + 2) = 396 values: two per thread
    constexpr int kSlice = (kTVY + 2) * (kTVX + 2);
    static_assert(kSlice <= 512, "two values per thread");
    float pre[2];                              // the next slice, in flight in registers
    auto fetch = [&](int z) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int k = threadIdx.x + 256 * r;
            const int hy = k / (kTVX + 2), hx = k - hy * (kTVX + 2);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1;
            float v = 0.f;
            if (k < kSlice && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                v = vol[(size_t)z * nxy + (size_t)gy * nx + gx];
            pre[r] = v;
        }
    };
    auto put = [&](int slot) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int k = threadIdx.x + 256 * r;
            if (k < kSlice) (&im[slot][0][0])[k] = pre[r];
        }
    };
    auto load_slice = [&](int z, int slot) {
        fetch(z);
        put(slot);
    };
    // variation of slice z (LDS slot z & 1; the image of slice z - 1 is in the other slot)
    auto var_slice = [&](int z) {
        const int s = z & 1, sm = s ^ 1;
        for (int k = threadIdx.x; k < (kTVY + 1) * (kTVX + 1); k += 256) {
            const int hy = k / (kTVX + 1), hx = k - hy * (kTVX + 1);
            const int gx = x0 + hx, gy = y0 + hy;
            if (gx >= nx || gy >= ny) continue;
            const float x = im[s][hy + 1][hx + 1];
            const float dx = gx > 0 ? x - im[s][hy + 1][hx] : 0.f;
            const float dy = gy > 0 ? x - im[s][hy][hx + 1] : 0.f;
            const float dz = z > 0 ? x - im[sm][hy + 1][hx + 1] : 0.f;
            va[s][hy][hx] = sqrtf(w.wx * dx * dx + w.wy * dy * dy + w.wz * dz * dz + w.eps);
        }
    };

    load_slice(zb - 1, (zb - 1) & 1);
    load_slice(zb, zb & 1);
    fetch(zb + 1);
    __syncthreads();
    var_slice(zb);
    float xm = im[(zb - 1) & 1][ty + 1][tx + 1];
    float xc = im[zb & 1][ty + 1][tx + 1];
    for (int z = zb; z < ze; z++) {
        const int s = z & 1, sp = s ^ 1;
        __syncthreads();                       // slot sp (slice z - 1) and var slot sp are free
        put(sp);                               // slice z + 1, fetched one iteration ago
        if (z + 1 < ze) fetch(z + 2);          // in flight during this slice's work
        const size_t i = base + (size_t)z * nxy + (size_t)iy * nx + ix;
        float bpi = 0.f, nmi = 0.f;
        if (UPDATE && own) {                   // issued before the barriers
            bpi = bp[i];
            nmi = norm[i];
        }
        __syncthreads();
        if (z + 1 < nz) var_slice(z + 1);
        __syncthreads();
        const float xp = im[sp][ty + 1][tx + 1];
        if (own) {
            const float v = va[s][ty][tx];
            float dx0 = ix > 0 ? xc - im[s][ty + 1][tx] : 0.f;
            float dy0 = iy > 0 ? xc - im[s][ty][tx + 1] : 0.f;
            float dz0 = z > 0 ? xc - xm : 0.f;
            float dx1 = 0.f, dy1 = 0.f, dz1 = 0.f, varx = v, vary = v, varz = v;
            if (ix < nx - 1) { dx1 = xc - im[s][ty + 1][tx + 2]; varx = va[s][ty][tx + 1]; }
            if (iy < ny - 1) { dy1 = xc - im[s][ty + 2][tx + 1]; vary = va[s][ty + 1][tx]; }
            if (z < nz - 1) { dz1 = xc - xp; varz = va[sp][ty][tx]; }
            dx0 *= w.wx; dx1 *= w.wx;
            dy0 *= w.wy; dy1 *= w.wy;
            dz0 *= w.wz; dz1 *= w.wz;
            const float a = (dx0 + dy0 + dz0) / v + dx1 / varx + dy1 / vary + dz1 / varz;
            const float c = (w.wx + w.wy + w.wz) / v + w.wx / varx + w.wy / vary + w.wz / varz;
            if (UPDATE) {
                o1[i] = xc - (bpi * acc + beta * a) / (nmi + beta * c);
            } else {
                o1[i] = a;
                o2[i] = c;
            }
            if (ovar) ovar[i] = v;
        }
        xm = xc;
        xc = xp;
    }
}

static dim3 tv_zgrid(int B, int nx, int ny, int nz, int& nzc) {
    nzc = ceil_div(nz, kTVZ);
    if ((long)nzc * B > 65535 || ceil_div(ny, kTVY) > 65535) fail("tv prior: ny / 4 and nz / 32 * batches must not exceed 65535");
    return dim3(ceil_div(nx, kTVX), ceil_div(ny, kTVY), nzc * B);
}

static dim3 tv_grid(int B, int nx, int ny, int nz) {
    if ((long)nz * B > 65535 || ny > 65535) fail("tv prior: ny and nz * batches must not exceed 65535");
    return dim3(ceil_div(nx, 64), ny, nz * B);
}

void tv_sqs(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz, int B, int nx, int ny,
            int nz, float eps, hipStream_t stream) {
    const TVW w{wx, wy, wz, eps};
    TimedLaunch t("tv_sqs", stream);
    if (tuning_get("tv_mode", 1) == 2) {
        tv_sqs_kernel<<<tv_grid(B, nx, ny, nz), 64, 0, stream>>>(s1, s2, var, img, nx, ny, nz, w);
    } else {
        int nzc = 1;
        const dim3 grid = tv_zgrid(B, nx, ny, nz, nzc);
        tv_zmarch_kernel<false><<<grid, 256, 0, stream>>>(s1, s2, var, img, nullptr, nullptr, 0.f, 0.f, nx, ny, nz,
                                                          nzc, w);
    }
    CTP_LAUNCH_CHECK();
}

void tv_sqs_update(float* out, float* var, const float* img, const float* bp, const float* norm, float beta, float acc,
                   float wx, float wy, float wz, int B, int nx, int ny, int nz, float eps, hipStream_t stream) {
    if (out == img) fail("tv_sqs_update: the output must not alias the input image (neighbours are read)");
    const TVW w{wx, wy, wz, eps};
    TimedLaunch t("tv_sqs_update", stream);
    if (tuning_get("tv_mode", 1) == 2) {
        tv_sqs_update_kernel<<<tv_grid(B, nx, ny, nz), 64, 0, stream>>>(out, var, img, bp, norm, beta, acc, nx, ny, nz,
                                                                        w);
    } else {
        int nzc = 1;
        const dim3 grid = tv_zgrid(B, nx, ny, nz, nzc);
        tv_zmarch_kernel<true><<<grid, 256, 0, stream>>>(out, nullptr, var, img, bp, norm, beta, acc, nx, ny, nz, nzc,
                                                         w);
    }
    CTP_LAUNCH_CHECK();
}

void tv_sqs_backward(float* dimg, const float* img, const float* g1, const float* g2, const float* gv, float wx,
                     float wy, float wz, int B, int nx, int ny, int nz, float eps, hipStream_t stream,
                     float* scratch) {
    const TVW w{wx, wy, wz, eps};
    const dim3 grid = tv_grid(B, nx, ny, nz);
    const size_t n = (size_t)B * nx * ny * nz;
    float2* pr = reinterpret_cast<float2*>(scratch);   // 2 floats per voxel, 8-byte aligned
    if (!pr) {
        Workspace& ws = workspace_for(stream);
        Plan pl;
        pl.add<float2>(n);
        ws.reserve(pl.bytes);
        pr = ws.take_n<float2>(n);
    }
    TimedLaunch t("tv_sqs_backward", stream);
    tv_grad_pr_kernel<<<grid, 64, 0, stream>>>(pr, img, g1, g2, gv, nx, ny, nz, w);
    CTP_LAUNCH_CHECK();
    tv_grad_x_kernel<<<grid, 64, 0, stream>>>(dimg, pr, img, g1, nx, ny, nz, w);
    CTP_LAUNCH_CHECK();
}

}  // namespace ctp
