// Siddon fan-beam projector (equiangular arc detector, stacked slices) on
// MI355X (gfx950): the reference's cuda/siddonFan.cu.
//
// Every detector row iv is a plane z = (iv - (nv-1)/2 + off_v) dv
// (cuda/siddonFan.cu:35-41): its rays are horizontal, so a ray runs inside one
// image slice (the reference's axis-parallel z rule: slice trunc of the row's
// index coordinate, no ray if it lies outside the grid) and all rows share the
// in-plane geometry of their column.
//
// Forward projection: one ray per lane, the traversal of the cone-beam
// projector (siddon_core.h) without the endpoint move the reference applies to
// cone beams only.  prj += line integral.
//
// Backprojection: the reference scatters with atomicAdd
// (cuda/siddon.cu:235-314); here it is voxel driven and deterministic.  For a
// pixel and a view, the candidate columns are those whose angular interval
// meets the pixel's (projection of the pixel corners); the chord of each
// candidate ray through the pixel square is the intersection of its two slab
// intervals from the source on (the segment the forward traversal assigns to
// that pixel), computed once and applied to every slice channel, whose
// projection values are the summed rows of that slice.  img += A^T prj.
#include "common.h"
#include "runtime.h"
#include "ops.h"
#include "siddon_core.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ctp {

// source and detector point of ray (iu, iv) of a view (cuda/siddonFan.cu:30-41)
__device__ inline void fan_ray(float deg, int iu, int iv, const Det& d, float dsd, float dso, float3& src, float3& dst) {
    const float c = cosf(deg), s = sinf(deg);
    const float a = (-((float)iu - (d.nu - 1) / 2.0f) - d.off_u) * d.du;
    const float z = ((float)iv - (d.nv - 1) / 2.0f + d.off_v) * d.dv;
    src = make_float3(dso * s, -dso * c, z);
    const float rx = dsd * sinf(a), ry = -dso + dsd * cosf(a);
    dst = make_float3(rx * c - ry * s, rx * s + ry * c, z);
}

// block: 64 columns x 4 rows of one view
__global__ __launch_bounds__(256) void siddon_fan_fp_kernel(float* __restrict__ prj, const float* __restrict__ img,
                                                            const float* __restrict__ deg, Grid g, Det d, int nview,
                                                            float dsd, float dso) {
    const int iu = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iv = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int view = blockIdx.z;
    if (iu >= d.nu || iv >= d.nv) return;
    const __amdgpu_buffer_rsrc_t vol =
        __builtin_amdgcn_make_buffer_rsrc((void*)img, (short)0, 4 * g.nx * g.ny * g.nz, 0x00020000);
    float3 s, t;
    fan_ray(deg[view], iu, iv, d, dsd, dso, s, t);
    SiddonRay r;
    if (!siddon_setup(r, s, t, g, false, false)) return;
    const float val = siddon_integral(vol, r);
    prj[((size_t)view * d.nv + iv) * d.nu + iu] += val;
}

// Per (view, column), FP64: 1/w_x, 1/w_y of the in-plane direction in index
// units (0 when the reference treats the axis as parallel: |w| <= 1e-6), the
// ray length |q - s| and the fixed index of a parallel axis.  The z rule is per
// row (host).
struct FanRay {
    float ix, iy;     // 1/w_x, 1/w_y
    float len;        // |dst - src| (mm); 0 when the ray misses the grid in-plane
    float fx, fy;     // fixed index of a parallel axis (as float)
    float sx, sy;     // source, index coordinates
    float pad;
};

__global__ __launch_bounds__(256) void siddon_fan_rays_kernel(FanRay* __restrict__ rays, const float* __restrict__ deg,
                                                              Grid g, Det d, int nview, float dsd, float dso) {
    const int iu = blockIdx.x * 256 + threadIdx.x, view = blockIdx.y;
    if (iu >= d.nu) return;
    const double c = cos((double)deg[view]), s = sin((double)deg[view]);
    const double a = (-((double)iu - (d.nu - 1) / 2.0) - d.off_u) * d.du;
    const double S[2] = {dso * s, -dso * c};
    const double rx = dsd * sin(a), ry = -dso + dsd * cos(a);
    const double Q[2] = {rx * c - ry * s, rx * s + ry * c};
    const double dd[2] = {g.dx, g.dy}, cc[2] = {g.cx, g.cy};
    const int nn[2] = {g.nx, g.ny};
    double si[2], w[2];
    bool par[2];
    for (int k = 0; k < 2; k++) {
        si[k] = (S[k] - cc[k]) / dd[k] + nn[k] / 2.0;
        w[k] = (Q[k] - S[k]) / dd[k];
        par[k] = !(fabs(w[k]) > 1e-6);
    }
    double amin = 0, amax = 1;
    bool hit = true;
    for (int k = 0; k < 2; k++) {
        if (!par[k]) {
            const double a0 = -si[k] / w[k], a1 = (nn[k] - si[k]) / w[k];
            amin = fmax(amin, fmin(a0, a1));
            amax = fmin(amax, fmax(a0, a1));
        } else if (si[k] < 0 || si[k] > nn[k]) {
            hit = false;
        }
    }
    if (!(amax > amin)) hit = false;
    FanRay R{};
    for (int k = 0; k < 2; k++) {
        const double fixed = par[k] ? (double)(int)(si[k] + amin * w[k]) : -1.0;
        if (par[k] && (fixed < 0 || fixed >= nn[k])) hit = false;
        (k == 0 ? R.fx : R.fy) = (float)fixed;
        (k == 0 ? R.ix : R.iy) = par[k] ? 0.f : (float)(1.0 / w[k]);
    }
    R.len = hit ? (float)sqrt((Q[0] - S[0]) * (Q[0] - S[0]) + (Q[1] - S[1]) * (Q[1] - S[1])) : 0.f;
    R.sx = (float)si[0];
    R.sy = (float)si[1];
    rays[(size_t)view * d.nu + iu] = R;
}

// column coordinate (column iu spans [iu, iu + 1)) of a physical in-plane point
__device__ inline float fan_ucoord(float px, float py, float c, float s, const Det& d, float dso) {
    // angle of the point seen from the source, relative to the central ray
    const float rx = px * c + py * s, ry = -px * s + py * c;
    const float a = atanf(rx / (ry + dso));
    return -(a / d.du + d.off_u) + (float)(d.nu - 1) / 2.0f + 0.5f;
}

template <int CH>
__global__ __launch_bounds__(256) void siddon_fan_bp_kernel(float* __restrict__ img, float* __restrict__ part,
                                                            const float* __restrict__ q, const FanRay* __restrict__ rays,
                                                            const float* __restrict__ deg, int C, int cs, int nchunk,
                                                            int vpg, Grid g, Det d, int nview, float dso) {
    const int ix = blockIdx.x * 16 + (threadIdx.x & 15);
    const int iy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int grp = blockIdx.z / nchunk;
    const int c0 = (blockIdx.z - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;
    const float xa = ((float)ix - g.nx / 2.f) * g.dx + g.cx, ya = ((float)iy - g.ny / 2.f) * g.dy + g.cy;
    const float big = 3.0e38f;
    float acc[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) acc[j] = 0.f;
    for (int view = vbeg; view < vend; view++) {
        const float c = cosf(deg[view]), s = sinf(deg[view]);
        float umin = 3.0e38f, umax = -3.0e38f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float u = fan_ucoord(xa + ((k & 1) ? g.dx : 0.f), ya + ((k & 2) ? g.dy : 0.f), c, s, d, dso);
            umin = fminf(umin, u);
            umax = fmaxf(umax, u);
        }
        const int k0 = max((int)floorf(umin) - 1, 0), k1 = min((int)floorf(umax) + 1, d.nu - 1);
        const FanRay* __restrict__ Rv = rays + (size_t)view * d.nu;
        const float* __restrict__ qv = q + (size_t)view * d.nu * cs + c0;
        for (int k = k0; k <= k1; k++) {
            const FanRay R = Rv[k];
            if (R.len == 0.f) continue;
            float x0, x1, y0, y1;
            if (R.ix != 0.f) {
                const float ta = ((float)ix - R.sx) * R.ix, tb = ((float)ix + 1.f - R.sx) * R.ix;
                x0 = fminf(ta, tb); x1 = fmaxf(ta, tb);
            } else {
                const bool in = (float)ix == R.fx;
                x0 = in ? -big : big; x1 = in ? big : -big;
            }
            if (R.iy != 0.f) {
                const float ta = ((float)iy - R.sy) * R.iy, tb = ((float)iy + 1.f - R.sy) * R.iy;
                y0 = fminf(ta, tb); y1 = fmaxf(ta, tb);
            } else {
                const bool in = (float)iy == R.fy;
                y0 = in ? -big : big; y1 = in ? big : -big;
            }
            const float w = fmaxf(fminf(x1, y1) - fmaxf(fmaxf(x0, y0), 0.f), 0.f) * R.len;
            if (w > 0.f) {
                const float* p = qv + (size_t)k * cs;
#pragma unroll
                for (int j = 0; j < CH; j++) acc[j] = fmaf(w, p[j], acc[j]);
            }
        }
    }
    store_bp_result<CH>(img, part, acc, grp, c0, C, (size_t)g.nx * g.ny, (size_t)iy * g.nx + ix);
}

void siddon_fan_fp(float* prj, const float* img, const float* deg_dev, int B, const Grid& g, const Det& d, int nview,
                   float dsd, float dso, hipStream_t stream) {
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    if (nimg >= (1ull << 29)) fail("fan Siddon projection: volumes of 2^29 voxels or more are not supported");
    if (d.nv > 65535 * 4) fail("fan Siddon projection: too many detector rows");
    for (int b = 0; b < B; b++) {
        TimedLaunch t("siddon_fan_fp", stream);
        siddon_fan_fp_kernel<<<dim3(ceil_div(d.nu, 64), ceil_div(d.nv, 4), nview), 256, 0, stream>>>(
            prj + b * nprj, img + b * nimg, deg_dev, g, d, nview, dsd, dso);
        CTP_LAUNCH_CHECK();
    }
}

void siddon_fan_bp(float* img, const float* prj, const float* deg_dev, int B, const Grid& g, const Det& d, int nview,
                   float dsd, float dso, hipStream_t stream) {
    (void)dsd;
    const int C = B * g.nz;
    int CH = 1;
    while (CH < C && CH < 32) CH *= 2;
    const int nchunk = ceil_div(C, CH);
    const int cs = nchunk * CH;
    // detector rows of each slice: the row's plane index coordinate truncated
    // (the reference's axis-parallel z rule), rows outside [0, nz) dropped
    std::vector<std::vector<ChanRowHost>> per(C);
    for (int iv = 0; iv < d.nv; iv++) {
        const double z = ((double)iv - (d.nv - 1) / 2.0 + d.off_v) * d.dv;
        const double zi = (z - g.cz) / g.dz + g.nz / 2.0;
        if (zi < 0 || zi > g.nz) continue;
        const int iz = (int)zi;
        if (iz < 0 || iz >= g.nz) continue;
        for (int b = 0; b < B; b++) per[b * g.nz + iz].push_back(ChanRowHost{iv, 1.f});
    }
    const size_t npix = (size_t)g.nx * g.ny;
    const int groups = bp_view_groups(npix, nchunk, nview);
    const int vpg = ceil_div(nview, groups);
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<FanRay>((size_t)nview * d.nu);
    pl.add<float>((size_t)nview * d.nu * cs);
    pl.add<int>(C + 1);
    size_t nrows = 0;
    for (auto& v : per) nrows += v.size();
    pl.add<ChanRowHost>(std::max<size_t>(nrows, 1));
    if (groups > 1) pl.add<float>((size_t)groups * C * npix);
    ws.reserve(pl.bytes);
    FanRay* rays = ws.take_n<FanRay>((size_t)nview * d.nu);
    float* q = ws.take_n<float>((size_t)nview * d.nu * cs);
    int* cstart = ws.take_n<int>(C + 1);
    ChanRowHost* crow = ws.take_n<ChanRowHost>(std::max<size_t>(nrows, 1));
    float* part = groups > 1 ? ws.take_n<float>((size_t)groups * C * npix) : nullptr;
    std::vector<int> cs_h(C + 1, 0);
    std::vector<ChanRowHost> cr_h;
    for (int c = 0; c < C; c++) {
        cs_h[c] = (int)cr_h.size();
        cr_h.insert(cr_h.end(), per[c].begin(), per[c].end());
    }
    cs_h[C] = (int)cr_h.size();
    CTP_CHECK(hipMemcpyAsync(cstart, cs_h.data(), sizeof(int) * (C + 1), hipMemcpyHostToDevice, stream));
    if (!cr_h.empty())
        CTP_CHECK(hipMemcpyAsync(crow, cr_h.data(), sizeof(ChanRowHost) * cr_h.size(), hipMemcpyHostToDevice, stream));
    // per batch item b the rows come from prj[b] (c = b * nz + z)
    rows_to_channel_last(q, prj, cstart, crow, C, cs, g.nz, d, nview, stream);
    TimedLaunch t("siddon_fan_bp", stream);
    siddon_fan_rays_kernel<<<dim3(ceil_div(d.nu, 256), nview), 256, 0, stream>>>(rays, deg_dev, g, d, nview, dsd, dso);
    CTP_LAUNCH_CHECK();
    dim3 grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), nchunk * groups);
#define CTP_SFB(N)                                                                                             \
    case N:                                                                                                    \
        siddon_fan_bp_kernel<N><<<grid, 256, 0, stream>>>(img, part, q, rays, deg_dev, C, cs, nchunk, vpg, g, d, \
                                                          nview, dso);                                        \
        break;
    switch (CH) {
        CTP_SFB(1) CTP_SFB(2) CTP_SFB(4) CTP_SFB(8) CTP_SFB(16) CTP_SFB(32)
        default: fail("fan Siddon bp: bad channel chunk");
    }
#undef CTP_SFB
    CTP_LAUNCH_CHECK();
    if (part) reduce_parts(img, part, groups, (size_t)C * npix, stream);
    CTP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace ctp
