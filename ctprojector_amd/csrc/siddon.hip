// Siddon cone-beam projector for arbitrary trajectories on MI355X (gfx950).
//
// Forward projection (cuda/siddonCone.cu:12-42, cuda/siddon.cu:80-233):
//   one ray per lane; a 256-thread block covers a 16x16 detector patch of one
//   view as four 8x8 wavefront tiles, so the 64 rays of a wavefront stay
//   spatially coherent and their voxel gathers share cache lines.  The three-way
//   axis choice of the reference's traversal is evaluated with selects, and the
//   next crossing of each axis is recomputed from its step count with an FMA
//   (alpha0 + n * dAlpha) instead of being accumulated.
//
// Backprojection (exact transpose of the forward projector, the reference's
// cuda/siddon.cu:235-314 uses atomicAdd): voxel driven and atomic free.  For
// each (voxel, view) the detector cells whose rays can cross the voxel lie in a
// window of at most WU x WV cells, found from the projective map of the view
// and a conservative bound of the voxel's shadow.  The chord of each candidate
// ray through the voxel is the length of the intersection of its three slab
// intervals (ray-box clipping), which is exactly the segment length the
// reference's traversal assigns to that voxel.  Ray data (1/direction per axis
// and projection value times ray length) come from a table prepared per call in
// FP64.  Every voxel is written by exactly one thread: results are
// deterministic.
#include "common.h"
#include "runtime.h"
#include "ops.h"
#include "siddon_core.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ctp {

// ---------------------------------------------------------------------------
// Forward projection
// ---------------------------------------------------------------------------

// Views whose central ray runs mostly along x read the volume transposed to
// [z][x][y], so that the memory-contiguous axis is the one across neighbouring
// detector columns (lanes) in every view; the traversal then runs on the
// swapped grid (Siddon's segment lengths do not depend on the axis order).
__global__ void siddon_swap_flags_kernel(int* __restrict__ flags, const float* __restrict__ pdc,
                                         const float* __restrict__ psrc, int nview) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nview) return;
    const float wx = pdc[3 * v] - psrc[3 * v], wy = pdc[3 * v + 1] - psrc[3 * v + 1];
    flags[v] = fabsf(wx) > fabsf(wy) ? 1 : 0;
}

// [z][y][x] -> [z][x][y]
__global__ __launch_bounds__(256) void transpose_xy_kernel(float* __restrict__ out, const float* __restrict__ in, int nx,
                                                           int ny) {
    __shared__ float tile[32][33];
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32, z = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const size_t plane = (size_t)nx * ny * z;
    for (int k = ty; k < 32; k += 8) {
        const int y = y0 + k, x = x0 + tx;
        tile[k][tx] = (x < nx && y < ny) ? in[plane + (size_t)y * nx + x] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int x = x0 + k, y = y0 + tx;
        if (x < nx && y < ny) out[plane + (size_t)x * ny + y] = tile[tx][k];
    }
}

// One 4-wavefront block of rays (a TU x 4TV detector tile) of one view.
// ACC: prj += line integral for rays that hit the grid (the reference's
// semantics, cuda/siddon.cu:230); otherwise prj = line integral (0 on a miss).
template <int TU, int TV, bool ACC>
__device__ inline void siddon_fp_block(float* __restrict__ prj, const float* __restrict__ img,
                                       const float* __restrict__ imgT, int swap, const float* __restrict__ pdc,
                                       const float* __restrict__ pdu, const float* __restrict__ pdv,
                                       const float* __restrict__ psrc, Grid g, const Det& d, int view, int ub, int vb) {
    static_assert(TU * TV == 64, "one wavefront per tile");
    // layout choice and the volume's buffer resource, built before any divergent
    // branch so that they stay in scalar registers (one view per block)
    const int sw = __builtin_amdgcn_readfirstlane(swap);
    const __amdgpu_buffer_rsrc_t vol = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(sw ? imgT : img), (short)0, 4 * g.nx * g.ny * g.nz, 0x00020000);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // a block is 4 tiles stacked along v
    const int iu = ub * TU + (lane % TU);
    const int iv = (vb * 4 + wave) * TV + (lane / TU);
    if (iu >= d.nu || iv >= d.nv) return;
    float* out = prj + ((size_t)view * d.nv + iv) * d.nu + iu;
    SiddonRay r;
    if (!siddon_setup(r, ld3(psrc, view), cell_centre(pdc, pdu, pdv, view, iu, iv, d), g, sw != 0)) {
        if (!ACC) *out = 0.f;
        return;
    }
    const float val = siddon_integral(vol, r);
    if (ACC) *out += val;
    else *out = val;
}

// Block -> (view, detector tile).  Workgroups are dispatched round-robin over
// the 8 XCDs (each with its own L2).  In XCD order, XCD x receives the x-th
// eighth of every view's tiles in row-major order, i.e. a band of detector rows
// whose rays cross one slab of the volume, and walks through the views with
// it; otherwise tiles are in plain view-major order.  b counts blocks from a
// multiple of 8.
__device__ inline void fp_tile_of(long b, int ntile, int xcd_order, int& view, int& tile) {
    const int per = (ntile + 7) >> 3;                 // tiles per XCD and view
    if (xcd_order) {
        const int xcd = (int)(b & 7);
        const long j = b >> 3;
        view = (int)(j / per);
        tile = xcd * per + (int)(j - (long)view * per);
    } else {
        view = (int)(b / (8L * per));
        tile = (int)(b - (long)view * 8 * per);
    }
}

template <int TU, int TV>
__global__ __launch_bounds__(256) void siddon_cone_fp_kernel(float* __restrict__ prj, const float* __restrict__ img,
                                                             const float* __restrict__ imgT, const int* __restrict__ swapf,
                                                             const float* __restrict__ pdc, const float* __restrict__ pdu,
                                                             const float* __restrict__ pdv, const float* __restrict__ psrc,
                                                             Grid g, Det d, int nview, int nub, int ntile,
                                                             int xcd_order) {
    int view, tile;
    fp_tile_of(blockIdx.x, ntile, xcd_order, view, tile);
    if (tile >= ntile || view >= nview) return;
    siddon_fp_block<TU, TV, true>(prj, img, imgT, swapf[view], pdc, pdu, pdv, psrc, g, d, view, tile % nub, tile / nub);
}

// ---------------------------------------------------------------------------
// Column-driven forward projection
// ---------------------------------------------------------------------------
// When det_v has no x or y component, the rays of one detector column lie in a
// vertical plane through the source: they share one line in the xy plane and
// differ only in their slope along z.  One wavefront takes one column.  The 2D
// traversal of the shared xy line (its cells and the interval of the in-plane
// distance s spent in each) is the same in every lane and is stepped in
// lockstep; each lane follows up to 5 rays (detector rows, as many as the
// column's rows that reach the grid need) through it, and batch items sharing
// the geometry are projected together (the traversal and the split of each
// cell done once for up to 4 volumes).  Inside
// an xy cell a ray's z index changes by less than 1 (steeper rays are traced
// ray by ray), so it meets z cells lo and lo + 1 at most: the s interval is
// split at the plane z = lo + 1, and both voxels come from one 8-byte load of
// the volume re-laid as z-contiguous columns, [y][x][z + 1] with P = nz + 3
// floats per column and zeros at z + 1 = 0, nz + 1, nz + 2.  The 3D chord is
// the s interval times sqrt(1 + slope^2): the sum is the line integral
// Siddon's traversal computes (cuda/siddon.cu:169-233), up to rounding.
//
// Rays that this form would not reproduce exactly are traced with the
// reference's traversal (siddon_setup / siddon_integral): a component of the
// moved segment within the axis-parallel tolerance (|d| <= 1e-6 in index
// units, except an exactly horizontal ray), and rays steeper than the bound.
// Columns whose xy line has an exactly zero x or y component, and views whose
// det_v is not along z, are traced ray by ray as a whole.

// [z][y][x] -> [y][x][z + 1] with zero padding (P floats per column); with
// nb > 1 volumes interleaved, element (y, x, zp) of volume b at
// ((y * nx + x) * P + zp) * nb + b
__global__ __launch_bounds__(256) void transpose_zcol_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                             int nx, int ny, int nz, int P, int nb = 1, int b = 0) {
    __shared__ float tile[32][33];
    const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int z = z0 + k, x = x0 + tx;
        tile[k][tx] = (x < nx && z < nz) ? in[((size_t)z * ny + y) * nx + x] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int x = x0 + k, z = z0 + tx;
        if (x >= nx) continue;
        float* col = out + ((size_t)y * nx + x) * P * nb + b;
        if (z < nz) col[(z + 1) * nb] = tile[tx][k];
        if (blockIdx.y == 0 && tx < 3) col[(tx == 0 ? 0 : nz + tx) * nb] = 0.f;
    }
}

// Ray constants of the column form for detector row iv: slope along z per mm
// of s (zs, index units), z index at s = 0 (z0), the 3D length per mm of s
// (sec), and whether the ray must be traced instead.
struct ColFPRay {
    float zs, z0, sec;
    bool exc;
};

__device__ inline ColFPRay colfp_ray(float3 S, float3 P, const Grid& g, float Lxy, float sO, float diag) {
    ColFPRay c;
    const float slope = (P.z - S.z) / Lxy;
    c.zs = slope / g.dz;
    c.z0 = (fmaf(slope, sO, S.z) - g.cz) / g.dz + g.nz * 0.5f;
    c.sec = sqrtf(fmaf(slope, slope, 1.f));
    float3 s = S, t = P;
    siddon_move(s, t, g);
    const float3 dd = siddon_index_dir(s, t, g);
    const bool okx = fabsf(dd.x) > kEps, oky = fabsf(dd.y) > kEps;
    const bool okz = fabsf(dd.z) > kEps || (dd.z == 0.f && P.z == S.z);
    c.exc = !(okx && oky && okz) || !(fabsf(c.zs) * diag <= 0.99f);
    return c;
}

// ACC: out += line integral if the ray hits; otherwise out = line integral (0 on a miss)
template <bool ACC>
__device__ inline void trace_into(float* out, __amdgpu_buffer_rsrc_t vol, float3 S, float3 P, Grid g) {
    SiddonRay r;
    if (siddon_setup(r, S, P, g, false)) {
        const float v = siddon_integral(vol, r);
        if (ACC) *out += v;
        else *out = v;
    } else if (!ACC) {
        *out = 0.f;
    }
}

// Rays base + 64 r + lane (r < R) of a detector column through its xy line:
// see fp_column.
struct ColLine {
    float s_in, sxn0, syn0, dsx, dsy, Lxy, sO, diag, nzf;
    int rx0, ry0, mxo, myo, off0;
};

// NB (1 or 2) volumes of one geometry at once: the traversal and the per-ray
// split of each cell are shared, only the accumulations are per volume
// (volume b at col + b * pstride); the volumes are interleaved in volz (z
// index zp of volume b at float zp * NB + b of a column), so the pairs of both
// volumes come in one 16-byte load (one 8-byte load for NB = 1).
template <int R, bool ACC, int NB>
__device__ inline void fp_column_chunk(float* __restrict__ col, __amdgpu_buffer_rsrc_t volz, const ColLine& L,
                                       float3 S, const float* __restrict__ pdc, const float* __restrict__ pdu,
                                       const float* __restrict__ pdv, const Grid& g, const Det& d, int view, int iu,
                                       int base, int hi_iv, long pstride) {
    const int lane = threadIdx.x & 63;
    const float s_in = L.s_in, sxn0 = L.sxn0, syn0 = L.syn0, dsx = L.dsx, dsy = L.dsy, nzf = L.nzf;
    const int rx0 = L.rx0, ry0 = L.ry0, mxo = L.mxo, myo = L.myo, off0 = L.off0;
    float zs[R], z0[R], rs[R], q1[R], fa[R], acc[R][NB], sec[R];
    bool up[R], act[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int iv = base + 64 * r + lane;
        ColFPRay c;
        c.exc = true;
        if (iv <= hi_iv) c = colfp_ray(S, cell_centre(pdc, pdu, pdv, view, iu, iv, d), g, L.Lxy, L.sO, L.diag);
        act[r] = iv <= hi_iv && !c.exc;
        // inactive lanes follow a flat ray below the grid (zero padding)
        zs[r] = act[r] ? c.zs : 0.f;
        z0[r] = act[r] ? c.z0 : -4.f;
        sec[r] = c.sec;
        rs[r] = zs[r] != 0.f ? 1.f / zs[r] : 0.f;
        q1[r] = zs[r] != 0.f ? (1.f - z0[r]) * rs[r] : 3.0e38f;   // s at z = lo + 1: lo * rs + q1
        up[r] = zs[r] >= 0.f;
        fa[r] = __builtin_amdgcn_fmed3f(floorf(fmaf(zs[r], s_in, z0[r])), -1.f, nzf);
#pragma unroll
        for (int b = 0; b < NB; b++) acc[r][b] = 0.f;
    }
    // min / max as medians with a sentinel (no canonicalisation of
    // loop-carried operands)
    constexpr float kBig = 3.0e38f;
    float sa = s_in, sxn = sxn0, syn = syn0;
    int off = off0, rx = rx0, ry = ry0;
    // one xy cell; the loop below runs it twice per trip so that the
    // carried values alternate registers instead of being copied
    auto cell = [&]() {
        // every lane holds the same values: the choice is made scalar
        const int bxi = __builtin_amdgcn_readfirstlane(sxn <= syn ? 1 : 0);
        const float sb = __builtin_amdgcn_fmed3f(bxi ? sxn : syn, sa, kBig);
        const float Lc = sb - sa;
        float lo[R], fb[R];
        float2 f[R][NB];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fb[r] = __builtin_amdgcn_fmed3f(floorf(fmaf(zs[r], sb, z0[r])), -1.f, nzf);
            lo[r] = __builtin_amdgcn_fmed3f(fa[r], fb[r], -kBig);
        }
        // the whole offset is the vector operand (checked against the
        // buffer's size); (lo + 1) >= 0 is the padded z index
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int o = 4 * NB * ((int)lo[r] + 1) + off;
            if constexpr (NB == 1) {
                f[r][0] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(volz, o, 0, 0));
            } else {
                static_assert(NB == 2, "1 or 2 volumes");
                // (v0[lo], v1[lo], v0[lo + 1], v1[lo + 1])
                const float4 q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(volz, o, 0, 0));
                f[r][0] = make_float2(q.x, q.z);
                f[r][1] = make_float2(q.y, q.w);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float tm = __builtin_amdgcn_fmed3f(fmaf(lo[r], rs[r], q1[r]), sa, sb);
            const float llo = fabsf(tm - (up[r] ? sa : sb));      // part below z = lo + 1
#pragma unroll
            for (int b = 0; b < NB; b++) acc[r][b] = fmaf(llo, f[r][b].x, fmaf(Lc - llo, f[r][b].y, acc[r][b]));
            fa[r] = fb[r];
        }
        sa = sb;
        off += bxi ? mxo : myo;
        rx -= bxi;
        ry -= bxi ^ 1;
        sxn = bxi ? sxn + dsx : sxn;
        syn = bxi ? syn : syn + dsy;
    };
    while ((rx | ry) >= 0) {
        cell();
        if ((rx | ry) < 0) break;
        cell();
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int iv = base + 64 * r + lane;
        if (act[r]) {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float* o = col + b * pstride + (size_t)iv * d.nu;
                if (ACC) *o += sec[r] * acc[r][b];
                else *o = sec[r] * acc[r][b];
            }
        }
    }
}

// rays per lane and traversal (a 320-row column is one traversal; four rays
// per lane measured 10.1 against 8.2 ms at 256^3/720)
template <int NB>
constexpr int col_rays_max() { return 5; }

template <bool ACC, int NB>
__device__ inline void fp_column(float* __restrict__ prj, __amdgpu_buffer_rsrc_t volz, const float* __restrict__ img,
                                 const float* __restrict__ pdc, const float* __restrict__ pdu,
                                 const float* __restrict__ pdv, const float* __restrict__ psrc, const Grid& g,
                                 const Det& d, int view, int iu, int P, long pstride, long istride) {
    const int lane = threadIdx.x & 63;
    float* col = prj + (size_t)view * d.nv * d.nu + iu;        // row iv at col[iv * nu]
    const float3 S = ld3(psrc, view);
    const float3 V = ld3(pdv, view);
    const float3 C0 = cell_centre(pdc, pdu, pdv, view, iu, 0, d);
    const float ex0 = C0.x - S.x, ey0 = C0.y - S.y;
    const float Lxy = sqrtf(ex0 * ex0 + ey0 * ey0);
    const float ex = ex0 / Lxy, ey = ey0 / Lxy;
    const float exi = ex / g.dx, eyi = ey / g.dy;
    const bool shared = V.x == 0.f && V.y == 0.f && Lxy > 0.f && exi != 0.f && eyi != 0.f;
    // the reference traversal for rays outside the column form, each volume
    auto trace_row = [&](int iv, float3 Pc) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const __amdgpu_buffer_rsrc_t vol = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(img + b * istride), (short)0, 4 * g.nx * g.ny * g.nz, 0x00020000);
            trace_into<ACC>(col + b * pstride + (size_t)iv * d.nu, vol, S, Pc, g);
        }
    };
    if (!shared) {
        for (int iv = lane; iv < d.nv; iv += 64) trace_row(iv, cell_centre(pdc, pdu, pdv, view, iu, iv, d));
        return;
    }
    // s = 0 at the foot of the perpendicular from the grid centre (keeps |s|
    // within the grid's half diagonal); (ox, oy): that point in index units
    const float sO = (g.cx - S.x) * ex + (g.cy - S.y) * ey;
    const float ox = (fmaf(ex, sO, S.x) - g.cx) / g.dx + g.nx * 0.5f;
    const float oy = (fmaf(ey, sO, S.y) - g.cy) / g.dy + g.ny * 0.5f;
    const float rxi = 1.f / exi, ryi = 1.f / eyi;
    const float sx0 = -ox * rxi, sx1 = ((float)g.nx - ox) * rxi;
    const float sy0 = -oy * ryi, sy1 = ((float)g.ny - oy) * ryi;
    const float s_in = fmaxf(fminf(sx0, sx1), fminf(sy0, sy1));
    const float s_out = fminf(fmaxf(sx0, sx1), fmaxf(sy0, sy1));
    const bool line_hits = s_in < s_out;
    const float diag = sqrtf(g.dx * g.dx + g.dy * g.dy);
    const float nzf = (float)g.nz;

    // rows whose rays reach the grid form one run [lo_iv, hi_iv]
    int lo_iv = d.nv, hi_iv = -1;
    bool any_exc = false;
    for (int k = 0; k * 64 < d.nv; k++) {
        const int iv = k * 64 + lane;
        bool hit = false, exc = false;
        if (iv < d.nv) {
            const ColFPRay c = colfp_ray(S, cell_centre(pdc, pdu, pdv, view, iu, iv, d), g, Lxy, sO, diag);
            const float za = fmaf(c.zs, s_in, c.z0), zb = fmaf(c.zs, s_out, c.z0);
            exc = c.exc;
            hit = !exc && line_hits && fmaxf(za, zb) > 0.f && fminf(za, zb) < nzf;
        }
        const unsigned long long mh = __ballot(hit), me = __ballot(exc);
        if (mh) {
            lo_iv = min(lo_iv, k * 64 + __ffsll((long long)mh) - 1);
            hi_iv = max(hi_iv, k * 64 + 63 - __clzll((long long)mh));
        }
        any_exc = any_exc || me != 0;
    }

    if (hi_iv >= lo_iv) {
        // 2D traversal start, the same in every lane (cuda/siddon.cu:113-149 in 2D)
        // (integer state made scalar: every lane computes the same values)
        int ix = __builtin_amdgcn_readfirstlane((int)floorf(fmaf(s_in, exi, ox)));
        int iy = __builtin_amdgcn_readfirstlane((int)floorf(fmaf(s_in, eyi, oy)));
        ix = min(max(ix, 0), g.nx - 1);
        iy = min(max(iy, 0), g.ny - 1);
        const bool px = __builtin_amdgcn_readfirstlane(exi > 0.f ? 1 : 0) != 0;
        const bool py = __builtin_amdgcn_readfirstlane(eyi > 0.f ? 1 : 0) != 0;
        const float sxn0 = ((float)(px ? ix + 1 : ix) - ox) * rxi;
        const float syn0 = ((float)(py ? iy + 1 : iy) - oy) * ryi;
        const float dsx = fabsf(rxi), dsy = fabsf(ryi);
        const int rx0 = px ? g.nx - 1 - ix : ix, ry0 = py ? g.ny - 1 - iy : iy;
        // byte offsets of the interleaved columns (NB volumes per z index)
        const int mxo = (px ? 4 : -4) * P * NB, myo = (py ? 4 : -4) * P * NB * g.nx;
        const int off0 = 4 * P * NB * (iy * g.nx + ix);   // column (ix, iy)
        const ColLine line{s_in, sxn0, syn0, dsx, dsy, Lxy, sO, diag, nzf, rx0, ry0, mxo, myo, off0};
        // up to col_rays_max rays per lane and traversal, as many as the rows need
        constexpr int RM = col_rays_max<NB>();
        for (int base = lo_iv; base <= hi_iv;) {
            const int nr = min(RM, (hi_iv - base + 64) / 64);
#define CTP_CHUNK(Rv) fp_column_chunk<Rv, ACC, NB>(col, volz, line, S, pdc, pdu, pdv, g, d, view, iu, base, hi_iv, \
                                                   pstride)
            if (nr == 1) CTP_CHUNK(1);
            else if (nr == 2) CTP_CHUNK(2);
            else if constexpr (RM >= 3) {
                if (nr == 3) CTP_CHUNK(3);
                else if constexpr (RM >= 4) {
                    if (nr == 4) CTP_CHUNK(4);
                    else if constexpr (RM >= 5) CTP_CHUNK(5);
                }
            }
#undef CTP_CHUNK
            base += 64 * nr;
        }
    }
    // traced rows, and (without accumulation) zeros for the rows that miss;
    // the rows written above are not touched again
    if (any_exc || !ACC) {
        for (int iv = lane; iv < d.nv; iv += 64) {
            const float3 Pc = cell_centre(pdc, pdu, pdv, view, iu, iv, d);
            if (colfp_ray(S, Pc, g, Lxy, sO, diag).exc) {
                trace_row(iv, Pc);
            } else if (!ACC && (iv < lo_iv || iv > hi_iv)) {
#pragma unroll
                for (int b = 0; b < NB; b++) col[b * pstride + (size_t)iv * d.nu] = 0.f;
            }
        }
    }
}

// block: 4 adjacent detector columns of one view, one per wavefront; NB
// volumes (consecutive batch items) per launch
template <int NB>
__global__ __launch_bounds__(256) void siddon_cone_fp_col_kernel(float* __restrict__ prj, const float* __restrict__ img,
                                                                 const float* __restrict__ imgz,
                                                                 const float* __restrict__ pdc,
                                                                 const float* __restrict__ pdu,
                                                                 const float* __restrict__ pdv,
                                                                 const float* __restrict__ psrc, Grid g, Det d,
                                                                 int nview, int ngroup, int P, int xcd_order,
                                                                 long pstride, long istride, int lvpb) {
    // a block is 4 >> lvpb adjacent columns of 1 << lvpb consecutive views
    int vg, grp;
    fp_tile_of(blockIdx.x, ngroup, xcd_order, vg, grp);
    const int w = threadIdx.x >> 6;
    const int iu = (grp << (2 - lvpb)) + (w & ((4 >> lvpb) - 1));
    const int view = (vg << lvpb) + (w >> (2 - lvpb));
    if (grp >= ngroup || view >= nview || iu >= d.nu) return;
    const __amdgpu_buffer_rsrc_t volz =
        __builtin_amdgcn_make_buffer_rsrc((void*)imgz, (short)0, 4 * NB * P * g.nx * g.ny, 0x00020000);
    fp_column<true, NB>(prj, volz, img, pdc, pdu, pdv, psrc, g, d, view, iu, P, pstride, istride);
}

// Block order of the forward projection: tuning cone_fp_order 1 plain, 2 XCD
// bands; by default plain for the column-driven form (mode 0; its work per
// detector column varies too much across a view for bands) and XCD bands for
// the ray-per-lane form.
static int fp_xcd_order(int mode) {
    const int o = tuning_get("cone_fp_order", 0);
    if (o < 0 || o > 2) fail("cone_fp_order must be 1 (plain) or 2 (XCD bands)");
    return o == 0 ? (mode == 0 ? 0 : 1) : (o == 2 ? 1 : 0);
}

static void siddon_cone_fp_grid(float* prj, const float* img, const ConeGeomDev& geo, int B, const Grid& g,
                                const Det& d, int nview, hipStream_t stream) {
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    if (nimg >= (1ull << 29)) fail("cone forward projection: volumes of 2^29 voxels or more are not supported");
    const int mode = tuning_get("cone_fp_mode", 0);           // 0: column form, 1: ray per lane
    if (mode != 0 && mode != 1) fail("cone_fp_mode must be 0 or 1");
    const int xcd_order = fp_xcd_order(mode);
    if (mode == 0) {
        const int P = g.nz + 3;
        const size_t ncol = (size_t)g.nx * g.ny * P;
        if (ncol >= (1ull << 29)) fail("cone forward projection: volumes of 2^29 voxels or more are not supported");
        // batch items sharing the geometry go through in pairs (byte offsets
        // within a pair's copies stay below 2^31); a kernel for four volumes
        // measured slower than two pairs (28.4 against 2 x 13.4 ms at
        // 256^3/720: five rays per lane need 142 VGPRs for four accumulators
        // each)
        int nbmax = 1;
        if (B >= 2 && 2 * ncol < (1ull << 29)) nbmax = 2;
        Workspace& ws = workspace_for(stream);
        Plan pl;
        pl.add<float>(ncol * nbmax);
        ws.reserve(pl.bytes);
        float* imgz = ws.take_n<float>(ncol * nbmax);
        // two views per block share more voxel columns in L1 than four adjacent
        // columns of one view (measured at 256^3/720: 8.03 ms against 8.3 ms)
        const int tv = tuning_get("cone_fp_views_per_block", 2);
        if (tv != 1 && tv != 2 && tv != 4) fail("cone_fp_views_per_block must be 1, 2 or 4");
        const int lvpb = tv == 4 ? 2 : tv == 2 ? 1 : 0;
        const int ngroup = ceil_div(d.nu, 4 >> lvpb);
        const long nblocks = (long)ceil_div(nview, 1 << lvpb) * 8 * ((ngroup + 7) / 8);
        if (nblocks > 0x7fffffffL) fail("cone forward projection: too many detector columns");
        for (int b = 0; b < B;) {
            const int nb = B - b >= 2 && nbmax >= 2 ? 2 : 1;
            {
                TimedLaunch t("siddon_transpose", stream);
                for (int k = 0; k < nb; k++) {
                    transpose_zcol_kernel<<<dim3(ceil_div(g.nx, 32), ceil_div(g.nz, 32), g.ny), 256, 0, stream>>>(
                        imgz, img + (b + k) * nimg, g.nx, g.ny, g.nz, P, nb, k);
                    CTP_LAUNCH_CHECK();
                }
            }
            TimedLaunch t("siddon_cone_fp", stream);
#define CTP_FPC(N)                                                                                                \
    siddon_cone_fp_col_kernel<N><<<(unsigned)nblocks, 256, 0, stream>>>(                                          \
        prj + b * nprj, img + b * nimg, imgz, geo.det_center, geo.det_u, geo.det_v, geo.src, g, d, nview, ngroup, P, \
        xcd_order, (long)nprj, (long)nimg, lvpb)
            if (nb == 2) CTP_FPC(2);
            else CTP_FPC(1);
#undef CTP_FPC
            CTP_LAUNCH_CHECK();
            b += nb;
        }
        return;
    }
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>(nimg);
    pl.add<int>(nview);
    ws.reserve(pl.bytes);
    float* imgT = ws.take_n<float>(nimg);
    int* swapf = ws.take_n<int>(nview);
    siddon_swap_flags_kernel<<<ceil_div(nview, 256), 256, 0, stream>>>(swapf, geo.det_center, geo.src, nview);
    CTP_LAUNCH_CHECK();
    const int TU = tuning_get("cone_fp_tile_u", 32);
    if (TU != 8 && TU != 16 && TU != 32 && TU != 64) fail("cone_fp_tile_u must be 8, 16, 32 or 64");
    const int TV = 64 / TU;
    const int nub = ceil_div(d.nu, TU), nvb = ceil_div(d.nv, 4 * TV);
    const int ntile = nub * nvb;
    const long nblocks = (long)nview * 8 * ((ntile + 7) / 8);
    if (nblocks > 0x7fffffffL) fail("cone forward projection: too many detector tiles");
    dim3 block(256), grid((unsigned)nblocks);
    for (int b = 0; b < B; b++) {
        {
            TimedLaunch t("siddon_transpose", stream);
            transpose_xy_kernel<<<dim3(ceil_div(g.nx, 32), ceil_div(g.ny, 32), g.nz), 256, 0, stream>>>(
                imgT, img + b * nimg, g.nx, g.ny);
            CTP_LAUNCH_CHECK();
        }
        TimedLaunch t("siddon_cone_fp", stream);
#define CTP_FP_T(A, Bv)                                                                                          \
    if (TU == A) siddon_cone_fp_kernel<A, Bv><<<grid, block, 0, stream>>>(prj + b * nprj, img + b * nimg, imgT, swapf, \
                                                                          geo.det_center, geo.det_u, geo.det_v,     \
                                                                          geo.src, g, d, nview, nub, ntile, \
                                                                          xcd_order);
        CTP_FP_T(8, 8) CTP_FP_T(16, 4) CTP_FP_T(32, 2) CTP_FP_T(64, 1)
#undef CTP_FP_T
        CTP_LAUNCH_CHECK();
    }
}

// ---------------------------------------------------------------------------
// Backprojection
// ---------------------------------------------------------------------------
// Per-view constants of the voxel-driven backprojector (64 bytes; read through
// the scalar cache since every lane of a block works on the same view).
struct alignas(16) ConeViewBP {
    float sx, sy, sz;       // source in image index space
    float hu;               // conservative half extent of a voxel shadow, detector cells (u)
    float pu[4];            // iu_f(X) numerator over (X, Y, Z, 1), X = voxel-centre index coordinates
    float pv[4];            // iv_f(X) numerator
    float pw[4];            // common denominator (> 0 in front of the source)
    float hv;               // half extent along v
    int flags;              // bit0: some ray of the view is parallel to an image axis
    float fixx, fixy;       // trunc(source index) per axis, for axis-parallel rays
    float fixz, pad0, pad1, pad2;
};

struct ConeBPPlan {
    std::vector<ConeViewBP> views;
    int wu = 0, wv = 0;
    int near_view = -1;     // a view whose source is inside or too close to the volume (no bounded window)
};

// Builds the per-view projective maps in FP64 from host geometry.  A view whose
// source is inside or too close to the volume has no bounded candidate window:
// it is reported in near_view, or refused when 'strict'.
static ConeBPPlan make_cone_bp_plan(const ConeGeomHost& h, const Grid& g, const Det& d, int nview,
                                    bool strict = true) {
    ConeBPPlan plan;
    plan.views.resize(nview);
    double wmax_u = 0, wmax_v = 0;
    for (int i = 0; i < nview; i++) {
        const double C[3] = {h.det_center[3 * i], h.det_center[3 * i + 1], h.det_center[3 * i + 2]};
        const double U[3] = {h.det_u[3 * i], h.det_u[3 * i + 1], h.det_u[3 * i + 2]};
        const double V[3] = {h.det_v[3 * i], h.det_v[3 * i + 1], h.det_v[3 * i + 2]};
        const double S[3] = {h.src[3 * i], h.src[3 * i + 1], h.src[3 * i + 2]};
        double n[3] = {U[1] * V[2] - U[2] * V[1], U[2] * V[0] - U[0] * V[2], U[0] * V[1] - U[1] * V[0]};
        const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        if (!(nn > 0)) fail("cone geometry: det_u and det_v are parallel in view " + std::to_string(i));
        double D = (C[0] - S[0]) * n[0] + (C[1] - S[1]) * n[1] + (C[2] - S[2]) * n[2];
        if (D < 0) { for (double& x : n) x = -x; D = -D; }
        if (!(D > 0)) fail("cone geometry: source lies in the detector plane in view " + std::to_string(i));
        // dual basis of (U, V) within the detector plane
        const double Us[3] = {(V[1] * n[2] - V[2] * n[1]) / nn, (V[2] * n[0] - V[0] * n[2]) / nn,
                              (V[0] * n[1] - V[1] * n[0]) / nn};
        const double Vs[3] = {(n[1] * U[2] - n[2] * U[1]) / nn, (n[2] * U[0] - n[0] * U[2]) / nn,
                              (n[0] * U[1] - n[1] * U[0]) / nn};
        // orientation: (U x V).n must be positive for the dual basis above
        double sgn = ((U[1] * V[2] - U[2] * V[1]) * n[0] + (U[2] * V[0] - U[0] * V[2]) * n[1] +
                      (U[0] * V[1] - U[1] * V[0]) * n[2]) > 0 ? 1.0 : -1.0;
        double Usx[3], Vsx[3];
        for (int k = 0; k < 3; k++) { Usx[k] = Us[k] * sgn; Vsx[k] = Vs[k] * sgn; }
        const double uS = (S[0] - C[0]) * Usx[0] + (S[1] - C[1]) * Usx[1] + (S[2] - C[2]) * Usx[2];
        const double vS = (S[0] - C[0]) * Vsx[0] + (S[1] - C[1]) * Vsx[1] + (S[2] - C[2]) * Vsx[2];
        // u_mm(x) = uS + D * ((x-S).Us) / ((x-S).n); cell index iu_f = u/du + off_u + (nu-1)/2
        const double ou = d.off_u + (d.nu - 1) / 2.0, ov = d.off_v + (d.nv - 1) / 2.0;
        // x_phys = (X - N/2) * dxyz + c with X in voxel index coordinates
        const double dd[3] = {g.dx, g.dy, g.dz};
        const double x0[3] = {g.cx - g.nx / 2.0 * g.dx, g.cy - g.ny / 2.0 * g.dy, g.cz - g.nz / 2.0 * g.dz};
        double Pw[4], Pu[4], Pv[4];
        Pw[3] = 0; Pu[3] = 0; Pv[3] = 0;
        for (int k = 0; k < 3; k++) {
            Pw[k] = n[k] * dd[k];
            Pw[3] += n[k] * (x0[k] - S[k]);
        }
        double qu[4], qv[4];   // (x-S).Us and (x-S).Vs as affine maps of X
        qu[3] = 0; qv[3] = 0;
        for (int k = 0; k < 3; k++) {
            qu[k] = Usx[k] * dd[k]; qu[3] += Usx[k] * (x0[k] - S[k]);
            qv[k] = Vsx[k] * dd[k]; qv[3] += Vsx[k] * (x0[k] - S[k]);
        }
        for (int k = 0; k < 4; k++) {
            Pu[k] = (uS / d.du + ou) * Pw[k] + D / d.du * qu[k];
            Pv[k] = (vS / d.dv + ov) * Pw[k] + D / d.dv * qv[k];
        }
        // bound over the volume: corners of [0,N]^3 in index coordinates
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300, wmin = 1e300;
        for (int cc = 0; cc < 8; cc++) {
            const double X[3] = {(cc & 1) ? (double)g.nx : 0.0, (cc & 2) ? (double)g.ny : 0.0, (cc & 4) ? (double)g.nz : 0.0};
            double w = Pw[3], a = Pu[3], b = Pv[3];
            for (int k = 0; k < 3; k++) { w += Pw[k] * X[k]; a += Pu[k] * X[k]; b += Pv[k] * X[k]; }
            wmin = std::min(wmin, w);
            if (w > 0) {
                umin = std::min(umin, a / w); umax = std::max(umax, a / w);
                vmin = std::min(vmin, b / w); vmax = std::max(vmax, b / w);
            }
        }
        const double bw = 0.5 * (std::fabs(Pw[0]) + std::fabs(Pw[1]) + std::fabs(Pw[2]));
        if (!(wmin - bw > 0.02 * D)) {
            if (strict)
                fail("cone backprojection: the source is inside or too close to the volume in view " +
                     std::to_string(i));
            plan.near_view = i;
            return plan;
        }
        auto hnum = [&](const double* P, double uc) {
            return 0.5 * (std::fabs(P[0] - uc * Pw[0]) + std::fabs(P[1] - uc * Pw[1]) + std::fabs(P[2] - uc * Pw[2]));
        };
        const double margin = 1e-3;
        const double hu = std::max(hnum(Pu, umin), hnum(Pu, umax)) / (wmin - bw) + margin;
        const double hv = std::max(hnum(Pv, vmin), hnum(Pv, vmax)) / (wmin - bw) + margin;
        wmax_u = std::max(wmax_u, 2 * hu);
        wmax_v = std::max(wmax_v, 2 * hv);

        ConeViewBP& cv = plan.views[i];
        cv.sx = (float)((S[0] - g.cx) / g.dx + g.nx / 2.0);
        cv.sy = (float)((S[1] - g.cy) / g.dy + g.ny / 2.0);
        cv.sz = (float)((S[2] - g.cz) / g.dz + g.nz / 2.0);
        // normalise the projective rows so that the denominator is O(1)
        const double scale = 1.0 / D;
        for (int k = 0; k < 4; k++) {
            cv.pu[k] = (float)(Pu[k] * scale);
            cv.pv[k] = (float)(Pv[k] * scale);
            cv.pw[k] = (float)(Pw[k] * scale);
        }
        cv.hu = (float)hu;
        cv.hv = (float)hv;
        cv.fixx = std::trunc((float)cv.sx);
        cv.fixy = std::trunc((float)cv.sy);
        cv.fixz = std::trunc((float)cv.sz);
        // axis-parallel rays: the direction component w_k(u, v) is affine over the
        // detector; flag the view when it can vanish on the detector.
        int flags = 0;
        const double umm[2] = {(0 - ou) * d.du, (d.nu - 1 - ou) * d.du};
        const double vmm[2] = {(0 - ov) * d.dv, (d.nv - 1 - ov) * d.dv};
        for (int k = 0; k < 3; k++) {
            double lo = 1e300, hi = -1e300, mag = 0;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    double w = C[k] + U[k] * umm[a] + V[k] * vmm[b] - S[k];
                    lo = std::min(lo, w); hi = std::max(hi, w); mag = std::max(mag, std::fabs(w));
                }
            if (lo <= 1e-5 * mag && hi >= -1e-5 * mag) flags |= 1;
        }
        cv.flags = flags;
        cv.pad0 = cv.pad1 = cv.pad2 = 0;
    }
    plan.wu = (int)std::floor(wmax_u) + 1;
    plan.wv = (int)std::floor(wmax_v) + 1;
    return plan;
}

// Ray table for one batch item: (1/w_x, 1/w_y, 1/w_z, p * |q - S|) per ray in
// image index space, w = q - S with q the detector cell centre.  Computed in
// FP64.  Rays whose segment [S, q] (after MoveSourceDstNearGrid) misses the
// grid get weight 0 (InitializeSiddon's no-intersection exits,
// cuda/siddon.cu:98-115,155-161); axis-parallel components (|d_k| <= 1e-6 in the
// reference, cuda/siddon.cu:15) are stored as 0.
__global__ __launch_bounds__(256) void siddon_bp_table_kernel(float4* __restrict__ table, const float* __restrict__ prj,
                                                              const float* __restrict__ pdc, const float* __restrict__ pdu,
                                                              const float* __restrict__ pdv, const float* __restrict__ psrc,
                                                              Grid g, Det d, int nview) {
    const long n = (long)d.nu * d.nv * nview;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int iu = (int)(i % d.nu);
    const int iv = (int)((i / d.nu) % d.nv);
    const int view = (int)(i / ((long)d.nu * d.nv));
    const double u = ((double)iu - d.off_u - (d.nu - 1) / 2.0) * d.du;
    const double v = ((double)iv - d.off_v - (d.nv - 1) / 2.0) * d.dv;
    double S[3], Q[3];
    for (int k = 0; k < 3; k++) {
        S[k] = psrc[3 * view + k];
        Q[k] = pdc[3 * view + k] + pdu[3 * view + k] * u + pdv[3 * view + k] * v;
    }
    const double dd[3] = {g.dx, g.dy, g.dz};
    const double cc[3] = {g.cx, g.cy, g.cz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    // MoveSourceDstNearGrid in FP64
    double Sm[3] = {S[0], S[1], S[2]}, Qm[3] = {Q[0], Q[1], Q[2]};
    {
        double ex = g.nx * (double)g.dx, ey = g.ny * (double)g.dy, ez = g.nz * (double)g.dz;
        double r = sqrt(ex * ex + ey * ey + ez * ez);
        double w[3] = {Q[0] - S[0], Q[1] - S[1], Q[2] - S[2]};
        double len = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(len < 1e-6)) for (int k = 0; k < 3; k++) w[k] /= len;
        double t0 = 0, t1 = 0;
        for (int k = 0; k < 3; k++) { t0 += (cc[k] - S[k]) * w[k]; t1 += (Q[k] - cc[k]) * w[k]; }
        if (t0 > r) for (int k = 0; k < 3; k++) Sm[k] = S[k] + (t0 - r) * w[k];
        if (t1 > r) for (int k = 0; k < 3; k++) Qm[k] = Q[k] - (t1 - r) * w[k];
    }
    double L = 0, si[3], wi[3], di[3];
    for (int k = 0; k < 3; k++) {
        L += (Q[k] - S[k]) * (Q[k] - S[k]);
        si[k] = (S[k] - cc[k]) / dd[k] + nn[k] / 2.0;
        wi[k] = (Q[k] - S[k]) / dd[k];
        di[k] = (Qm[k] - Sm[k]) / dd[k];
    }
    L = sqrt(L);
    // segment [Sm, Qm] against the grid box [0, n]^3
    const double smi[3] = {(Sm[0] - cc[0]) / dd[0] + nn[0] / 2.0, (Sm[1] - cc[1]) / dd[1] + nn[1] / 2.0,
                           (Sm[2] - cc[2]) / dd[2] + nn[2] / 2.0};
    double amin = 0, amax = 1;
    bool hit = true;
    float inv[3];
    for (int k = 0; k < 3; k++) {
        if (fabs(di[k]) > 1e-6) {
            double a0 = (0 - smi[k]) / di[k], a1 = (nn[k] - smi[k]) / di[k];
            amin = fmax(amin, fmin(a0, a1));
            amax = fmin(amax, fmax(a0, a1));
        } else if (smi[k] < 0 || smi[k] > nn[k]) {
            hit = false;
        }
        inv[k] = (fabs(di[k]) > 1e-6) ? (float)(1.0 / wi[k]) : 0.f;
    }
    if (!(amax > amin)) hit = false;
    for (int k = 0; k < 3; k++) {
        if (!(fabs(di[k]) > 1e-6)) {
            // the reference's first voxel along an axis-parallel direction is
            // trunc(coordinate); outside [0, n) the ray contributes nothing
            int fixed = (int)(smi[k] + amin * di[k]);
            if (fixed < 0 || fixed >= nn[k]) hit = false;
        }
    }
    (void)si;
    const float p = prj[i];
    table[i] = make_float4(inv[0], inv[1], inv[2], hit ? (float)(p * L) : 0.f);
}

// Chord of one candidate ray through the voxel at (ex, ey, ez) + [-1/2, 1/2]^3
// relative to the source (index units): intersection of the three slab
// intervals t_k = (X_k - S_k) * inv_k, clipped to t >= 0 (the ray starts at the
// source), times the ray weight (projection value x ray length).
template <bool SLOW>
__device__ inline float chord(float4 r, float ex, float ey, float ez, bool inx, bool iny, bool inz) {
    const float mx = ex * r.x, my = ey * r.y, mz = ez * r.z;
    float x0 = fmaf(fabsf(r.x), -0.5f, mx), x1 = fmaf(fabsf(r.x), 0.5f, mx);
    float y0 = fmaf(fabsf(r.y), -0.5f, my), y1 = fmaf(fabsf(r.y), 0.5f, my);
    float z0 = fmaf(fabsf(r.z), -0.5f, mz), z1 = fmaf(fabsf(r.z), 0.5f, mz);
    if (SLOW) {
        // axis-parallel component (stored as 0): inside the slab iff the voxel
        // index equals the ray's constant index along that axis
        const float big = 3.0e38f;
        if (r.x == 0.f) { x0 = inx ? -big : big; x1 = inx ? big : -big; }
        if (r.y == 0.f) { y0 = iny ? -big : big; y1 = iny ? big : -big; }
        if (r.z == 0.f) { z0 = inz ? -big : big; z1 = inz ? big : -big; }
    }
    const float tin = fmaxf(fmaxf(fmaxf(x0, y0), z0), 0.f);
    const float tout = fminf(fminf(x1, y1), z1);
    return fmaxf(tout - tin, 0.f) * r.w;
}

template <int WU, int WV, int NZR, bool SLOW>
__device__ inline void bp_view(float* acc, const ConeViewBP& cv, const float4* __restrict__ tab, int nu, int nv,
                               int ix, int iy, int iz0, float xc, float yc) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)tab, (short)0, 16 * nu * nv, 0x00020000);
    const float ex = xc - cv.sx, ey = yc - cv.sy;
    const float bu = fmaf(cv.pu[0], xc, fmaf(cv.pu[1], yc, cv.pu[3]));
    const float bv = fmaf(cv.pv[0], xc, fmaf(cv.pv[1], yc, cv.pv[3]));
    const float bw = fmaf(cv.pw[0], xc, fmaf(cv.pw[1], yc, cv.pw[3]));
    const bool inx = (float)ix == cv.fixx, iny = (float)iy == cv.fixy;
#pragma unroll
    for (int k = 0; k < NZR; k++) {
        const float zc = (float)(iz0 + k) + 0.5f;
        const float rd = __builtin_amdgcn_rcpf(fmaf(cv.pw[2], zc, bw));
        const float uc = fmaf(cv.pu[2], zc, bu) * rd;
        const float vc = fmaf(cv.pv[2], zc, bv) * rd;
        const int u0 = min(max((int)ceilf(uc - cv.hu), 0), nu - WU);
        const int v0 = min(max((int)ceilf(vc - cv.hv), 0), nv - WV);
        const float ez = zc - cv.sz;
        const bool inz = (float)(iz0 + k) == cv.fixz;
        const int base = 16 * (v0 * nu + u0);
        float4 r[WV][WU];
#pragma unroll
        for (int jv = 0; jv < WV; jv++)
#pragma unroll
            for (int ju = 0; ju < WU; ju++) {
                const int o = base + 16 * (jv * nu + ju);
                r[jv][ju] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0));
            }
        float a = 0.f;
#pragma unroll
        for (int jv = 0; jv < WV; jv++)
#pragma unroll
            for (int ju = 0; ju < WU; ju++) a += chord<SLOW>(r[jv][ju], ex, ey, ez, inx, iny, inz);
        acc[k] += a;
    }
}

template <int WU, int WV, int NZR>
__global__ __launch_bounds__(256) void siddon_cone_bp_kernel(float* __restrict__ img, const float4* __restrict__ table,
                                                             const ConeViewBP* __restrict__ views, int nview, Grid g,
                                                             int nu, int nv) {
    const int ix = blockIdx.x * 16 + (threadIdx.x & 15);
    const int iy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int iz0 = blockIdx.z * NZR;
    if (ix >= g.nx || iy >= g.ny) return;
    const float xc = ix + 0.5f, yc = iy + 0.5f;
    float acc[NZR];
#pragma unroll
    for (int k = 0; k < NZR; k++) acc[k] = 0.f;
    const size_t nuv = (size_t)nu * nv;
    for (int view = 0; view < nview; view++) {
        const ConeViewBP cv = views[view];
        const float4* tab = table + view * nuv;
        if (cv.flags & 1) bp_view<WU, WV, NZR, true>(acc, cv, tab, nu, nv, ix, iy, iz0, xc, yc);
        else bp_view<WU, WV, NZR, false>(acc, cv, tab, nu, nv, ix, iy, iz0, xc, yc);
    }
#pragma unroll
    for (int k = 0; k < NZR; k++) {
        const int iz = iz0 + k;
        if (iz < g.nz) img[((size_t)iz * g.ny + iy) * g.nx + ix] += acc[k];
    }
}

// ---------------------------------------------------------------------------
// Column mode: det_v parallel to z and det_u orthogonal to z in every view (the
// circular / helical scanners of the reference's examples).  A ray's direction
// then has x, y components that depend on its detector column only and a z
// component that depends on its row only, so the x/y slab interval of a
// candidate column is shared by every voxel of a z-run and every candidate
// row.  Tables: per ray p*|q-S| (4 B), per view and column (1/w_x, 1/w_y), per
// view and row 1/w_z.
//
// The reference treats a direction component as axis-parallel when its
// index-space extent after MoveSourceDstNearGrid is at most 1e-6
// (cuda/siddon.cu:15,98-115).  Here a column (row) component is parallel when
// it is exactly zero; a ray for which the two rules disagree (0 < |d_k| <= 1e-6)
// gets table weight 0 and is traced and scattered by the table kernel itself,
// with the forward projector's arithmetic, so the result is the reference's
// in every case.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_zero_i32_kernel(int* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Weight table of the column-mode backprojector, laid out so that a voxel's
// candidate window is contiguous: entry (view, iv, iu, jv) = weight of ray
// (iv + jv, iu) for jv < dup (the window height; 0 past the last row).  One
// thread computes the dup weights of an entry and stores them together.

// Ray i = (view, iv, iu) in FP64 with the reference's rules: returns the weight
// (0 for rays that miss, or that follow the axis-parallel rule on an axis they
// are not parallel to); with 'scatter', adds those exception rays to img by
// the exact traversal.
__device__ inline float coltable_exact(long i, int iu, int iv, int view, bool scatter, float* __restrict__ img,
                                      const float* __restrict__ prj, const float* __restrict__ pdc,
                                      const float* __restrict__ pdu, const float* __restrict__ pdv,
                                      const float* __restrict__ psrc, const Grid& g, const Det& d) {
    const double u = ((double)iu - d.off_u - (d.nu - 1) / 2.0) * d.du;
    const double v = ((double)iv - d.off_v - (d.nv - 1) / 2.0) * d.dv;
    double S[3], Q[3];
    for (int k = 0; k < 3; k++) {
        S[k] = psrc[3 * view + k];
        Q[k] = pdc[3 * view + k] + pdu[3 * view + k] * u + pdv[3 * view + k] * v;
    }
    const double dd[3] = {g.dx, g.dy, g.dz};
    const double cc[3] = {g.cx, g.cy, g.cz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    double Sm[3] = {S[0], S[1], S[2]}, Qm[3] = {Q[0], Q[1], Q[2]};
    {
        double ex = g.nx * (double)g.dx, ey = g.ny * (double)g.dy, ez = g.nz * (double)g.dz;
        double r = sqrt(ex * ex + ey * ey + ez * ez);
        double w[3] = {Q[0] - S[0], Q[1] - S[1], Q[2] - S[2]};
        double len = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(len < 1e-6)) for (int k = 0; k < 3; k++) w[k] /= len;
        double t0 = 0, t1 = 0;
        for (int k = 0; k < 3; k++) { t0 += (cc[k] - S[k]) * w[k]; t1 += (Q[k] - cc[k]) * w[k]; }
        if (t0 > r) for (int k = 0; k < 3; k++) Sm[k] = S[k] + (t0 - r) * w[k];
        if (t1 > r) for (int k = 0; k < 3; k++) Qm[k] = Q[k] - (t1 - r) * w[k];
    }
    double L = 0, wi[3], di[3], smi[3];
    bool rpar[3], cpar[3];
    for (int k = 0; k < 3; k++) {
        L += (Q[k] - S[k]) * (Q[k] - S[k]);
        wi[k] = (Q[k] - S[k]) / dd[k];
        di[k] = (Qm[k] - Sm[k]) / dd[k];
        smi[k] = (Sm[k] - cc[k]) / dd[k] + nn[k] / 2.0;
        rpar[k] = !(fabs(di[k]) > 1e-6);
        cpar[k] = (Q[k] - S[k]) == 0.0;
    }
    L = sqrt(L);
    double amin = 0, amax = 1;
    bool hit = true;
    for (int k = 0; k < 3; k++) {
        if (!rpar[k]) {
            double a0 = (0 - smi[k]) / di[k], a1 = (nn[k] - smi[k]) / di[k];
            amin = fmax(amin, fmin(a0, a1));
            amax = fmin(amax, fmax(a0, a1));
        } else if (smi[k] < 0 || smi[k] > nn[k]) {
            hit = false;
        }
    }
    if (!(amax > amin)) hit = false;
    for (int k = 0; k < 3; k++) {
        if (rpar[k]) {
            int fixed = (int)(smi[k] + amin * di[k]);
            if (fixed < 0 || fixed >= nn[k]) hit = false;
        }
    }
    const bool exception = rpar[0] != cpar[0] || rpar[1] != cpar[1] || rpar[2] != cpar[2];
    const float p = prj[i];
    if (scatter && exception && p != 0.f) {
        SiddonRay r;
        Grid gg = g;
        if (siddon_setup(r, ld3(psrc, view), cell_centre(pdc, pdu, pdv, view, iu, iv, d), gg, false)) {
            bool inside = true;
            while (inside) {
                const int off = r.off;
                const float len = siddon_step(r, inside);
                atomicAdd(img + (off >> 2), p * len);
            }
        }
    }
    return (hit && !exception) ? (float)(p * L) : 0.f;
}

// Column and row tables of the column-mode BP (1/w_x, 1/w_y per column from
// the ray of row 0; 1/w_z, |1/w_z|/2 per row from the ray of column 0; FP64,
// 0 for an axis the ray is exactly parallel to) and the view's parallel-ray
// flag.  One thread per column and per row of every view.
__global__ __launch_bounds__(256) void siddon_bp_colaxes_kernel(float2* __restrict__ utab, float2* __restrict__ vtab,
                                                                float2* __restrict__ colw, float* __restrict__ roww,
                                                                ConeViewBP* __restrict__ views,
                                                                const float* __restrict__ pdc,
                                                                const float* __restrict__ pdu,
                                                                const float* __restrict__ pdv,
                                                                const float* __restrict__ psrc, Grid g, Det d,
                                                                int nview) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = d.nu + d.nv;
    if (t >= (long)per * nview) return;
    const int view = (int)(t / per), k = (int)(t % per);
    const bool column = k < d.nu;
    const int iu = column ? k : 0, iv = column ? 0 : k - d.nu;
    const double u = ((double)iu - d.off_u - (d.nu - 1) / 2.0) * d.du;
    const double v = ((double)iv - d.off_v - (d.nv - 1) / 2.0) * d.dv;
    const double dd[3] = {g.dx, g.dy, g.dz};
    double w[3], wi[3];
    bool cpar[3];
    for (int a = 0; a < 3; a++) {
        const double S = psrc[3 * view + a];
        const double Q = pdc[3 * view + a] + pdu[3 * view + a] * u + pdv[3 * view + a] * v;
        w[a] = Q - S;
        wi[a] = (Q - S) / dd[a];
        cpar[a] = (Q - S) == 0.0;
    }
    // (in column geometry Q - S has x, y from the column and z from the row)
    if (column) {
        utab[(long)view * d.nu + iu] =
            make_float2(cpar[0] ? 0.f : (float)(1.0 / wi[0]), cpar[1] ? 0.f : (float)(1.0 / wi[1]));
        colw[(long)view * d.nu + iu] = make_float2((float)w[0], (float)w[1]);
        if (cpar[0] || cpar[1]) views[view].flags = 1;
    } else {
        const float q = cpar[2] ? 0.f : (float)(1.0 / wi[2]);
        vtab[(long)view * d.nv + iv] = make_float2(q, 0.5f * fabsf(q));
        roww[(long)view * d.nv + iv] = (float)w[2];
        if (cpar[2]) views[view].flags = 1;
    }
}

// Column-mode BP weights per call.  A ray with every component of its (moved)
// direction at least 1e-5 voxels from zero is regular under the reference's
// rules whatever the rounding: its weight is prj * |Q - S|, with hits decided
// by the geometry of the chords themselves (a ray that misses the grid has no
// chord in any voxel).  Rays near an axis direction take the FP64 path.
// Fast-path weight of ray (view, iv, iu) (w = prj * |Q - S|), or -1 when the
// ray is not regular and needs the FP64 path.
__device__ inline float coltable_fast(long i, int iu, int iv, int view, const float* __restrict__ prj,
                                      const float2* __restrict__ colw, const float* __restrict__ roww,
                                      const float* __restrict__ psrc, const Grid& g, const Det& d) {
    // Q - S (the axes kernel's FP64 values rounded to FP32: zero exactly when zero)
    const float2 cw = colw[(long)view * d.nu + iu];
    const float3 wf = make_float3(cw.x, cw.y, roww[(long)view * d.nv + iv]);
    // fraction of the ray kept by the move towards the grid (it scales the
    // components, which are compared with a 10x margin: approximate
    // reciprocals are enough)
    const float3 Sf = ld3(psrc, view);
    const float L = __builtin_amdgcn_sqrtf(wf.x * wf.x + wf.y * wf.y + wf.z * wf.z);
    const float rL = __builtin_amdgcn_rcpf(L);
    const float ex = (float)g.nx * g.dx, ey = (float)g.ny * g.dy, ez = (float)g.nz * g.dz;
    const float rr = sqrtf(ex * ex + ey * ey + ez * ez);
    const float t0 = ((g.cx - Sf.x) * wf.x + (g.cy - Sf.y) * wf.y + (g.cz - Sf.z) * wf.z) * rL;
    const float frac = (L - fmaxf(t0 - rr, 0.f) - fmaxf(L - t0 - rr, 0.f)) * rL;
    const bool regular = fabsf(wf.x) * frac > 1e-5f * g.dx && fabsf(wf.y) * frac > 1e-5f * g.dy &&
                         fabsf(wf.z) * frac > 1e-5f * g.dz && frac > 0.f;
    return regular ? prj[i] * L : -1.f;
}

// One thread per column and 8 consecutive rows of a view (grid: column
// blocks, row groups, views): the weights of rays iv0 .. iv0 + 8 + dup - 2
// (dup <= 4) fill the entries of rows iv0 .. iv0 + 7.  Rays that are not
// regular get weight 0 here and flag their row for siddon_bp_colpatch_kernel.
constexpr int kTabRows = 8;

template <int DUP>
__global__ __launch_bounds__(256) void siddon_bp_coltable_kernel(float* __restrict__ wtab, int* __restrict__ rowirr,
                                                                 const float* __restrict__ prj,
                                                                 const float2* __restrict__ colw,
                                                                 const float* __restrict__ roww,
                                                                 const float* __restrict__ psrc, Grid g, Det d) {
    const int iu = blockIdx.x * 256 + threadIdx.x, iv0 = blockIdx.y * kTabRows, view = blockIdx.z;
    if (iu >= d.nu) return;
    const long i0 = ((long)view * d.nv + iv0) * d.nu + iu;
    const int nrow = min(kTabRows + DUP - 1, d.nv - iv0);     // rays of this thread
    float wr[kTabRows + DUP - 1];
#pragma unroll
    for (int r = 0; r < kTabRows + DUP - 1; r++) {
        float w = 0.f;
        if (r < nrow) {
            w = coltable_fast(i0 + (long)r * d.nu, iu, iv0 + r, view, prj, colw, roww, psrc, g, d);
            if (w < 0.f && r < kTabRows) rowirr[(long)view * d.nv + iv0 + r] = 1;
        }
        wr[r] = fmaxf(w, 0.f);
    }
#pragma unroll
    for (int r = 0; r < kTabRows; r++) {
        if (r < nrow) {
            float* out = wtab + (i0 + (long)r * d.nu) * DUP;
            if constexpr (DUP == 2) {
                *reinterpret_cast<float2*>(out) = make_float2(wr[r], wr[r + 1]);
            } else {
#pragma unroll
                for (int jv = 0; jv < DUP; jv++) out[jv] = wr[r + jv];
            }
        }
    }
}

// The rays left at weight 0 by siddon_bp_coltable_kernel: FP64 weight into
// their dup entries, and the exact scatter of exception rays.  Grid: (row
// groups of 8, views); blocks of rows without a flag return at once.
__global__ __launch_bounds__(256) void siddon_bp_colpatch_kernel(float* __restrict__ wtab, float* __restrict__ img,
                                                                 const int* __restrict__ rowirr,
                                                                 const float* __restrict__ prj,
                                                                 const float2* __restrict__ colw,
                                                                 const float* __restrict__ roww,
                                                                 const float* __restrict__ pdc,
                                                                 const float* __restrict__ pdu,
                                                                 const float* __restrict__ pdv,
                                                                 const float* __restrict__ psrc, Grid g, Det d,
                                                                 int dup) {
    const int view = blockIdx.y;
    for (int r = 0; r < kTabRows; r++) {
        const int iv = blockIdx.x * kTabRows + r;
        if (iv >= d.nv) return;
        if (rowirr[(long)view * d.nv + iv] == 0) continue;
        for (int iu = threadIdx.x; iu < d.nu; iu += 256) {
            const long i = ((long)view * d.nv + iv) * d.nu + iu;
            if (coltable_fast(i, iu, iv, view, prj, colw, roww, psrc, g, d) >= 0.f) continue;
            const float w = coltable_exact(i, iu, iv, view, true, img, prj, pdc, pdu, pdv, psrc, g, d);
            for (int jv = 0; jv < dup && jv <= iv; jv++) wtab[(i - (long)jv * d.nu) * dup + jv] = w;
        }
    }
}

// N consecutive floats (N <= 4) from a buffer with one load instruction
template <int N>
__device__ inline void load_floats(float* out, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
    static_assert(N >= 1 && N <= 4, "1 to 4 floats");
    if constexpr (N == 1) {
        out[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else if constexpr (N == 2) {
        const float2 v = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
        out[0] = v.x; out[1] = v.y;
    } else if constexpr (N == 3) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
        out[0] = __uint_as_float(v[0]); out[1] = __uint_as_float(v[1]); out[2] = __uint_as_float(v[2]);
    } else {
        const float4 v = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
        out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    }
}

// WU x WV is the largest candidate window compiled in; wu x wv (<= WU x WV,
// uniform across the block) is the window used.
// NB volumes of one geometry: wt[b] is volume b's weight table of the view;
// the chords are computed once and accumulated into acc[k * NB + b].
template <int WU, int WV, int NZR, bool SLOW, int NB, bool AHEAD = true>
__device__ __forceinline__ void bp_view_col(float* acc, const ConeViewBP& cv, const float* const* wt,
                                   const float2* __restrict__ ut, const float2* __restrict__ vt, int nu, int nv, int ix,
                                   int iy, int iz0, float xc, float yc, int wu, int wv) {
    __amdgpu_buffer_rsrc_t rw[NB];
#pragma unroll
    for (int b = 0; b < NB; b++)
        rw[b] = __builtin_amdgcn_make_buffer_rsrc((void*)wt[b], (short)0, 4 * (nu * nv * WV + 16), 0x00020000);
    const __amdgpu_buffer_rsrc_t ru = __builtin_amdgcn_make_buffer_rsrc((void*)ut, (short)0, 8 * nu, 0x00020000);
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc((void*)vt, (short)0, 8 * (nv + 4), 0x00020000);
    const float ex = xc - cv.sx, ey = yc - cv.sy;
    // pu[2] = pw[2] = 0 in column mode: the u coordinate and the denominator
    // do not depend on z
    const float rd = __builtin_amdgcn_rcpf(fmaf(cv.pw[0], xc, fmaf(cv.pw[1], yc, cv.pw[3])));
    const float uc = fmaf(cv.pu[0], xc, fmaf(cv.pu[1], yc, cv.pu[3])) * rd;
    const float bv = fmaf(cv.pv[0], xc, fmaf(cv.pv[1], yc, cv.pv[3]));
    // Candidate rows of voxel k start at v0 = floor(vc_k - hv + 1), clamped.
    // Rays of a nonzero chord lie strictly within hv - 1e-3 of vc (the margin
    // of the host's bound) and wv > 2 hv, so every origin between
    // ceil(vc - hv) and floor(vc - hv + 1) holds all of them: vc - hv + 1 is
    // evaluated as an affine function of k (a few ulp from the direct form),
    // and an origin moved by that rounding only adds or drops rows of zero
    // chord, which leave the sums unchanged.
    const float y1 = fmaf(cv.pv[2], (float)iz0 + 0.5f, bv) * rd + (1.f - cv.hv), dy1 = cv.pv[2] * rd;
    // window origins clamped as floats (one median) before the conversion
    const int u0 = (int)__builtin_amdgcn_fmed3f(ceilf(uc - cv.hu), 0.f, (float)(nu - wu));
    const int u0w = 4 * WV * u0;   // byte offset of the window's first column in a table row
    const float v0max = (float)(nv - wv);
    const bool inx = (float)ix == cv.fixx, iny = (float)iy == cv.fixy;
    const float big = 3.0e38f;
    float lo[WU], hi[WU];
#pragma unroll
    for (int j = 0; j < WU; j++) {
        lo[j] = 0.f;
        hi[j] = 0.f;
        if (j >= wu) continue;
        const float2 q = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(ru, 8 * (u0 + j), 0, 0));
        const float mx = ex * q.x, my = ey * q.y;
        float x0 = fmaf(fabsf(q.x), -0.5f, mx), x1 = fmaf(fabsf(q.x), 0.5f, mx);
        float y0 = fmaf(fabsf(q.y), -0.5f, my), y1 = fmaf(fabsf(q.y), 0.5f, my);
        if (SLOW) {
            if (q.x == 0.f) { x0 = inx ? -big : big; x1 = inx ? big : -big; }
            if (q.y == 0.f) { y0 = iny ? -big : big; y1 = iny ? big : -big; }
        }
        lo[j] = fmaxf(fmaxf(x0, y0), 0.f);
        hi[j] = fmaxf(fminf(x1, y1), lo[j]);       // empty interval: hi = lo
    }
    // voxel k's candidate rows (row table) and window weights
    auto fetch = [&](int k, float* qv, float (*span)[WU * WV]) {
        // origin >= 0 after the clamp, so the conversion's truncation is the floor
        const int v0 = (int)__builtin_amdgcn_fmed3f(k == 0 ? y1 : fmaf((float)k, dy1, y1), 0.f, v0max);
        // The weights of the whole WU x WV window are contiguous in the window
        // layout (entry (iv, iu, jv) = weight of ray (iv + jv, iu)): one to
        // four loads; the row table takes one load per two rows.  Entries past
        // (wu, wv) are read but not used: a window load reaches at most
        // (WU - wu) * WV floats past the view's last entry and a row-table load
        // at most one row past nv - 1 (both tables are padded).
        // one 24-bit multiply-add (4 WV nu < 2^24: nu < 2^20 is checked on the host)
        const int wbase = (int)__umul24((unsigned)v0, (unsigned)(4 * WV * nu)) + u0w, vbase = 8 * v0;
#pragma unroll
        for (int jv = 0; jv < WV; jv += 2) {
            const int r = min(jv, wv - 1);
            if (jv + 1 < WV) load_floats<4>(qv + 2 * jv, rv, vbase, 8 * r);
            else load_floats<2>(qv + 2 * jv, rv, vbase, 8 * r);
        }
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int c = 0; c < WU * WV; c += 4) {
                constexpr int N = WU * WV;
                if (N - c >= 4) load_floats<4>(span[b] + c, rw[b], wbase, 4 * c);
                else if (N - c == 3) load_floats<3>(span[b] + c, rw[b], wbase, 4 * c);
                else if (N - c == 2) load_floats<2>(span[b] + c, rw[b], wbase, 4 * c);
                else load_floats<1>(span[b] + c, rw[b], wbase, 4 * c);
            }
    };
    // chords of voxel k's candidate rays, weighted, into its accumulators
    auto apply = [&](int k, const float* qv, const float (*span)[WU * WV]) {
        const float zc = (float)(iz0 + k) + 0.5f;
        const float ez = zc - cv.sz;
        const bool inz = (float)(iz0 + k) == cv.fixz;
        float2 qz[WV];
#pragma unroll
        for (int jv = 0; jv < WV; jv++) qz[jv] = make_float2(qv[2 * jv], qv[2 * jv + 1]);
        float a[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) a[b] = 0.f;
#pragma unroll
        for (int jv = 0; jv < WV; jv++) {
            if (jv >= wv) continue;
            // z-slab interval of the ray: ez * qz -+ |qz| / 2
            float z0 = fmaf(ez, qz[jv].x, -qz[jv].y), z1 = fmaf(ez, qz[jv].x, qz[jv].y);
            if (SLOW && qz[jv].x == 0.f) { z0 = inz ? -big : big; z1 = big; }
            // overlap of [z0, z1] with [lo, hi] (lo <= hi): both ends clamped into it
#pragma unroll
            for (int j = 0; j < WU; j++) {
                if (j >= wu) continue;
                const float ch = __builtin_amdgcn_fmed3f(z1, lo[j], hi[j]) - __builtin_amdgcn_fmed3f(z0, lo[j], hi[j]);
#pragma unroll
                for (int b = 0; b < NB; b++) a[b] = fmaf(ch, span[b][j * WV + jv], a[b]);
            }
        }
#pragma unroll
        for (int b = 0; b < NB; b++) acc[k * NB + b] += a[b];
    };
    // The loads do not depend on the chords, so (with AHEAD) they are issued
    // ahead of them and their latencies overlap the arithmetic: the whole
    // z-run's loads before the first chord when they take at most 32 floats
    // per lane, else voxel k + 1's loads before voxel k's chords when two
    // voxels' take at most 32, else voxel by voxel.
    constexpr int per = 2 * WV + NB * WU * WV;   // floats loaded per voxel
    if constexpr (AHEAD && NZR * per <= 32) {
        float qv[NZR][2 * WV];
        float span[NZR][NB][WU * WV];
#pragma unroll
        for (int k = 0; k < NZR; k++) fetch(k, qv[k], span[k]);
#pragma unroll
        for (int k = 0; k < NZR; k++) apply(k, qv[k], span[k]);
    } else if constexpr (AHEAD && 2 * per <= 32) {
        float qv[2][2 * WV];
        float span[2][NB][WU * WV];
        fetch(0, qv[0], span[0]);
#pragma unroll
        for (int k = 0; k < NZR; k++) {
            if (k + 1 < NZR) fetch(k + 1, qv[(k + 1) & 1], span[(k + 1) & 1]);
            apply(k, qv[k & 1], span[k & 1]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NZR; k++) {
            float qv[2 * WV];
            float span[NB][WU * WV];
            fetch(k, qv, span);
            apply(k, qv, span);
        }
    }
}

// weight tables (per-call window layout) of up to 4 volumes of one geometry
struct WTabs {
    const float* p[4];
};

// One block of the column-mode backprojector: 16x16 voxel columns x NZR slices,
// img += A^T y over all views, for NB volumes (volume b at img + b * istride).
// AHEAD: loads issued ahead of the chords (bp_view_col).
template <int WU, int WV, int NZR, int NB, bool AHEAD = true>
__device__ __forceinline__ void bp_col_block(float* __restrict__ img, long istride, const WTabs& wtab,
                                    const float2* __restrict__ utab, const float2* __restrict__ vtab,
                                    const ConeViewBP* __restrict__ views, int nview, const Grid& g, int nu, int nv,
                                    int bx, int by, int bz, int wu = WU, int wv = WV) {
    const int ix = bx * 16 + (threadIdx.x & 15);
    const int iy = by * 16 + (threadIdx.x >> 4);
    const int iz0 = bz * NZR;
    if (ix >= g.nx || iy >= g.ny) return;
    const float xc = ix + 0.5f, yc = iy + 0.5f;
    float acc[NZR * NB];
#pragma unroll
    for (int k = 0; k < NZR * NB; k++) acc[k] = 0.f;
    const size_t nuv = (size_t)nu * nv;
    for (int view = 0; view < nview; view++) {
        const ConeViewBP cv = const_load(views, view);   // scalar loads (read-only table)
        const float* wt[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) wt[b] = wtab.p[b] + view * nuv * WV;
        const float2* ut = utab + (size_t)view * nu;
        const float2* vt = vtab + (size_t)view * nv;
        if (cv.flags & 1)
            bp_view_col<WU, WV, NZR, true, NB, AHEAD>(acc, cv, wt, ut, vt, nu, nv, ix, iy, iz0, xc, yc, wu, wv);
        else bp_view_col<WU, WV, NZR, false, NB, AHEAD>(acc, cv, wt, ut, vt, nu, nv, ix, iy, iz0, xc, yc, wu, wv);
    }
#pragma unroll
    for (int k = 0; k < NZR; k++) {
        const int iz = iz0 + k;
        if (iz < g.nz)
#pragma unroll
            for (int b = 0; b < NB; b++) img[b * istride + ((size_t)iz * g.ny + iy) * g.nx + ix] += acc[k * NB + b];
    }
}

template <int WU, int WV, int NZR, int NB>
__global__ __launch_bounds__(256) void siddon_cone_bp_col_kernel(float* __restrict__ img, long istride, WTabs wtab,
                                                                 const float2* __restrict__ utab,
                                                                 const float2* __restrict__ vtab,
                                                                 const ConeViewBP* __restrict__ views, int nview,
                                                                 Grid g, int nu, int nv) {
    bp_col_block<WU, WV, NZR, NB>(img, istride, wtab, utab, vtab, views, nview, g, nu, nv, blockIdx.x, blockIdx.y,
                                  blockIdx.z);
}

template <int WU, int WV, int NZR, int NB>
static void launch_bp_col_n(float* img, long istride, const WTabs& wtab, const float2* utab, const float2* vtab,
                            const ConeViewBP* views, int nview, const Grid& g, const Det& d, hipStream_t stream) {
    dim3 block(256), grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), ceil_div(g.nz, NZR));
    TimedLaunch t("siddon_cone_bp", stream);
    siddon_cone_bp_col_kernel<WU, WV, NZR, NB><<<grid, block, 0, stream>>>(img, istride, wtab, utab, vtab, views, nview,
                                                                           g, d.nu, d.nv);
    CTP_LAUNCH_CHECK();
}

// nb volumes (1, 2 or 4; 2 and 4 with the default z-run only)
template <int WU, int WV>
static void launch_bp_col(float* img, long istride, int nb, const WTabs& wtab, const float2* utab,
                          const float2* vtab, const ConeViewBP* views, int nview, const Grid& g, const Det& d,
                          hipStream_t stream) {
    // default: 8 voxels per lane for one volume with a 2x2 window (12.15
    // against 12.46 ms at 256^3 / 720 views with the loads issued ahead;
    // 3x3: 13.49 against 13.48 ms), else 4
    const int nzr = tuning_get("cone_bp_zrun", nb == 1 && WU == 2 && WV == 2 ? 8 : 4);
    if (nb == 4 && nzr == 4) return launch_bp_col_n<WU, WV, 4, 4>(img, istride, wtab, utab, vtab, views, nview, g, d, stream);
    if (nb == 2 && nzr == 4) return launch_bp_col_n<WU, WV, 4, 2>(img, istride, wtab, utab, vtab, views, nview, g, d, stream);
    if (nb != 1) fail("cone backprojection: several volumes per launch need cone_bp_zrun 4");
    switch (nzr) {
        case 1: launch_bp_col_n<WU, WV, 1, 1>(img, istride, wtab, utab, vtab, views, nview, g, d, stream); break;
        case 2: launch_bp_col_n<WU, WV, 2, 1>(img, istride, wtab, utab, vtab, views, nview, g, d, stream); break;
        case 4: launch_bp_col_n<WU, WV, 4, 1>(img, istride, wtab, utab, vtab, views, nview, g, d, stream); break;
        case 8: launch_bp_col_n<WU, WV, 8, 1>(img, istride, wtab, utab, vtab, views, nview, g, d, stream); break;
        default: fail("cone_bp_zrun must be 1, 2, 4 or 8");
    }
}

// every view has det_v = (0, 0, v) and det_u = (ux, uy, 0) exactly, and the
// FP32 projective rows have no z term in u and in the denominator
static bool column_mode(const ConeGeomHost& h, const ConeBPPlan& plan, int nview) {
    for (int i = 0; i < nview; i++) {
        if (h.det_v[3 * i] != 0.f || h.det_v[3 * i + 1] != 0.f || h.det_u[3 * i + 2] != 0.f) return false;
        const ConeViewBP& cv = plan.views[i];
        if (cv.pu[2] != 0.f || cv.pw[2] != 0.f) return false;
    }
    return true;
}

template <int WU, int WV, int NZR>
static void launch_bp_n(float* img, const float4* table, const ConeViewBP* views, int nview, const Grid& g,
                        const Det& d, hipStream_t stream) {
    dim3 block(256), grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), ceil_div(g.nz, NZR));
    TimedLaunch t("siddon_cone_bp", stream);
    siddon_cone_bp_kernel<WU, WV, NZR><<<grid, block, 0, stream>>>(img, table, views, nview, g, d.nu, d.nv);
    CTP_LAUNCH_CHECK();
}

template <int WU, int WV>
static void launch_bp(float* img, const float4* table, const ConeViewBP* views, int nview, const Grid& g, const Det& d,
                      hipStream_t stream) {
    const int nzr = tuning_get("cone_bp_zrun", 4);
    switch (nzr) {
        case 1: launch_bp_n<WU, WV, 1>(img, table, views, nview, g, d, stream); break;
        case 2: launch_bp_n<WU, WV, 2>(img, table, views, nview, g, d, stream); break;
        case 4: launch_bp_n<WU, WV, 4>(img, table, views, nview, g, d, stream); break;
        case 8: launch_bp_n<WU, WV, 8>(img, table, views, nview, g, d, stream); break;
        default: fail("cone_bp_zrun must be 1, 2, 4 or 8");
    }
}

// ---------------------------------------------------------------------------
// Ray-driven backprojection: the reference's transpose (cuda/siddon.cu:235-314,
// cuda/siddonCone.cu:44-74), one lane per ray, the forward projector's
// traversal (siddon_setup / siddon_step), every segment added to its voxel
// with an atomic add.  It takes every geometry, and serves the ones the
// voxel-driven window backprojector does not: a voxel shadow wider than 4
// detector cells (fine detectors, strong magnification) and sources inside or
// close to the volume.  The summation order of the atomics is not fixed, as in
// the reference.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void siddon_cone_bp_scatter_kernel(float* __restrict__ img,
                                                                     const float* __restrict__ prj,
                                                                     const float* __restrict__ pdc,
                                                                     const float* __restrict__ pdu,
                                                                     const float* __restrict__ pdv,
                                                                     const float* __restrict__ psrc, Grid g, Det d,
                                                                     int nview) {
    const long n = (long)d.nu * d.nv * nview;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p = prj[i];
    if (p == 0.f) return;                  // adds nothing
    const int iu = (int)(i % d.nu);
    const int iv = (int)((i / d.nu) % d.nv);
    const int view = (int)(i / ((long)d.nu * d.nv));
    const float3 S = make_float3(psrc[3 * view], psrc[3 * view + 1], psrc[3 * view + 2]);
    const float u = ((float)iu - d.off_u - (float)(d.nu - 1) / 2.0f) * d.du;
    const float v = ((float)iv - d.off_v - (float)(d.nv - 1) / 2.0f) * d.dv;
    const float3 T = make_float3(pdc[3 * view] + pdu[3 * view] * u + pdv[3 * view] * v,
                                 pdc[3 * view + 1] + pdu[3 * view + 1] * u + pdv[3 * view + 1] * v,
                                 pdc[3 * view + 2] + pdu[3 * view + 2] * u + pdv[3 * view + 2] * v);
    SiddonRay r;
    if (!siddon_setup(r, S, T, g, false)) return;
    char* base = reinterpret_cast<char*>(img);
    // The traversal of siddon_step, with every crossing position computed as
    // first crossing + crossings so far x spacing instead of by repeated
    // addition: the voxel sequence is the same, and a segment's rounding no
    // longer grows with the number of steps (the reference's accumulated
    // crossings drift by up to ~1% on grids of a thousand voxels or more).
    const float ax0 = r.ax, ay0 = r.ay, az0 = r.az;
    float a = r.a;
    int kx = 0, ky = 0, kz = 0;
    int off = r.off, rx = r.rx, ry = r.ry, rz = r.rz;
    for (;;) {
        const float cx = fmaf((float)kx, r.dax, ax0), cy = fmaf((float)ky, r.day, ay0);
        const float cz = fmaf((float)kz, r.daz, az0);
        const float anow = fminf(fminf(cx, cy), cz);
        const bool bx = cx == anow;
        const bool by = !bx && cy == anow;
        const bool bz = !bx && !by;
        atomicAdd(reinterpret_cast<float*>(base + off), p * (anow - a));
        a = anow;
        off += bx ? r.mx : (by ? r.my : r.mz);
        rx -= bx ? 1 : 0;
        ry -= by ? 1 : 0;
        rz -= bz ? 1 : 0;
        kx += bx ? 1 : 0;
        ky += by ? 1 : 0;
        kz += bz ? 1 : 0;
        if ((rx | ry | rz) < 0) break;
    }
}

static void siddon_cone_bp_scatter(float* img, const float* prj, const ConeGeomDev& geo, int B, const Grid& g,
                                   const Det& d, int nview, hipStream_t stream) {
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    for (int b = 0; b < B; b++) {
        TimedLaunch t("siddon_cone_bp", stream);
        siddon_cone_bp_scatter_kernel<<<(unsigned)ceil_div((long)nprj, 256), 256, 0, stream>>>(
            img + b * nimg, prj + b * nprj, geo.det_center, geo.det_u, geo.det_v, geo.src, g, d, nview);
        CTP_LAUNCH_CHECK();
    }
}

static void siddon_cone_bp_grid(float* img, const float* prj, const ConeGeomDev& geo, const ConeGeomHost& hgeo,
                                int B, const Grid& g, const Det& d, int nview, hipStream_t stream) {
    // cone_bp_mode: 0 = window backprojector where the geometry allows it (column
    // mode when det_v is along z), 1 = general window mode, 2 = ray-driven scatter
    const int bp_mode = tuning_get("cone_bp_mode", 0);
    if (bp_mode < 0 || bp_mode > 2) fail("cone_bp_mode must be 0, 1 or 2");
    if ((size_t)d.nu * d.nv * nview / 256 >= 0x7fffffffULL) fail("cone backprojection: too many rays per call");
    ConeBPPlan plan = make_cone_bp_plan(hgeo, g, d, nview, false);
    if (bp_mode == 2 || plan.near_view >= 0 || plan.wu > 4 || plan.wv > 4) {
        siddon_cone_bp_scatter(img, prj, geo, B, g, d, nview, stream);
        return;
    }
    int wu = plan.wu, wv = plan.wv;
    wu = std::min(std::max(wu, 1), d.nu);
    wv = std::min(std::max(wv, 1), d.nv);
    const bool col = bp_mode == 0 && column_mode(hgeo, plan, nview) && d.nu < (1 << 20);
    const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    // column mode: batch items sharing the geometry are backprojected up to 4
    // per launch (default z-run; their weight tables within 4 GiB together)
    const size_t tab = nprj * wv + 16;              // window layout, padded: window loads
    int nbmax = 1;
    if (col && tuning_get("cone_bp_zrun", 4) == 4)
        while (nbmax < 4 && 2 * nbmax <= B && (size_t)(2 * nbmax) * tab * 4 <= (4ull << 30)) nbmax *= 2;
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<ConeViewBP>(nview);
    if (col) {
        pl.add<float>(tab * nbmax);
        pl.add<float2>((size_t)nview * d.nu);
        pl.add<float2>((size_t)nview * d.nv + 4);
        pl.add<float2>((size_t)nview * d.nu);
        pl.add<float>((size_t)nview * d.nv);
        pl.add<int>((size_t)nview * d.nv);
    } else {
        pl.add<float4>(nprj);
    }
    ws.reserve(pl.bytes);
    ConeViewBP* dviews = ws.take_n<ConeViewBP>(nview);
    float* wtab = nullptr;
    float2* utab = nullptr;
    float2* vtab = nullptr;
    float2* colw = nullptr;
    float* roww = nullptr;
    int* rowirr = nullptr;
    float4* table = nullptr;
    if (col) {
        wtab = ws.take_n<float>(tab * nbmax);
        utab = ws.take_n<float2>((size_t)nview * d.nu);
        vtab = ws.take_n<float2>((size_t)nview * d.nv + 4);
        colw = ws.take_n<float2>((size_t)nview * d.nu);
        roww = ws.take_n<float>((size_t)nview * d.nv);
        rowirr = ws.take_n<int>((size_t)nview * d.nv);
    } else {
        table = ws.take_n<float4>(nprj);
    }
    // column mode: the table kernel flags the views that need the parallel-ray
    // handling (a zero table entry); the geometric flag is conservative
    if (col)
        for (ConeViewBP& cv : plan.views) cv.flags = 0;
    CTP_CHECK(hipMemcpyAsync(dviews, plan.views.data(), sizeof(ConeViewBP) * nview, hipMemcpyHostToDevice, stream));
    if (col) {
        if (d.nv > 65535 || nview > 65535) fail("cone backprojection: more than 65535 rows or views");
        {
            TimedLaunch tt("siddon_bp_table", stream);
            siddon_bp_colaxes_kernel<<<ceil_div((long)nview * (d.nu + d.nv), 256), 256, 0, stream>>>(
                utab, vtab, colw, roww, dviews, geo.det_center, geo.det_u, geo.det_v, geo.src, g, d, nview);
            CTP_LAUNCH_CHECK();
        }
        for (int b = 0; b < B;) {
            const int nb = B - b >= 4 && nbmax >= 4 ? 4 : (B - b >= 2 && nbmax >= 2 ? 2 : 1);
            WTabs wt{{nullptr, nullptr, nullptr, nullptr}};
            for (int k = 0; k < nb; k++) {
                TimedLaunch tt("siddon_bp_table", stream);
                float* wk = wtab + k * tab;
                const float* pk = prj + (b + k) * nprj;
                wt.p[k] = wk;
                fill_zero_i32_kernel<<<ceil_div((long)nview * d.nv, 256), 256, 0, stream>>>(rowirr,
                                                                                             (size_t)nview * d.nv);
                CTP_LAUNCH_CHECK();
                const dim3 tg(ceil_div(d.nu, 256), ceil_div(d.nv, kTabRows), nview);
                switch (wv) {
                    case 1: siddon_bp_coltable_kernel<1><<<tg, 256, 0, stream>>>(wk, rowirr, pk, colw, roww, geo.src, g, d); break;
                    case 2: siddon_bp_coltable_kernel<2><<<tg, 256, 0, stream>>>(wk, rowirr, pk, colw, roww, geo.src, g, d); break;
                    case 3: siddon_bp_coltable_kernel<3><<<tg, 256, 0, stream>>>(wk, rowirr, pk, colw, roww, geo.src, g, d); break;
                    default: siddon_bp_coltable_kernel<4><<<tg, 256, 0, stream>>>(wk, rowirr, pk, colw, roww, geo.src, g, d); break;
                }
                CTP_LAUNCH_CHECK();
                siddon_bp_colpatch_kernel<<<dim3(ceil_div(d.nv, kTabRows), nview), 256, 0, stream>>>(
                    wk, img + (b + k) * nimg, rowirr, pk, colw, roww, geo.det_center, geo.det_u, geo.det_v, geo.src, g,
                    d, wv);
                CTP_LAUNCH_CHECK();
            }
            float* im = img + b * nimg;
            b += nb;
#define CTP_BPC_CASE(A, Bv)                                                                                    \
    if (wu == A && wv == Bv) {                                                                                 \
        launch_bp_col<A, Bv>(im, (long)nimg, nb, wt, utab, vtab, dviews, nview, g, d, stream);                 \
        continue;                                                                                              \
    }
            CTP_BPC_CASE(1, 1) CTP_BPC_CASE(1, 2) CTP_BPC_CASE(2, 1) CTP_BPC_CASE(2, 2)
            CTP_BPC_CASE(2, 3) CTP_BPC_CASE(3, 2) CTP_BPC_CASE(3, 3) CTP_BPC_CASE(1, 3) CTP_BPC_CASE(3, 1)
            CTP_BPC_CASE(1, 4) CTP_BPC_CASE(4, 1) CTP_BPC_CASE(2, 4) CTP_BPC_CASE(4, 2) CTP_BPC_CASE(3, 4)
            CTP_BPC_CASE(4, 3) CTP_BPC_CASE(4, 4)
#undef CTP_BPC_CASE
            fail("cone backprojection: unsupported window");
        }
        // host plan vector must outlive the async copy
        CTP_CHECK(hipStreamSynchronize(stream));
        return;
    }
    for (int b = 0; b < B; b++) {
        float* im = img + b * nimg;
        {
            TimedLaunch tt("siddon_bp_table", stream);
            siddon_bp_table_kernel<<<ceil_div((long)nprj, 256), 256, 0, stream>>>(table, prj + b * nprj, geo.det_center,
                                                                                geo.det_u, geo.det_v, geo.src, g, d, nview);
            CTP_LAUNCH_CHECK();
        }
#define CTP_BP_CASE(A, Bv) \
    if (wu == A && wv == Bv) { launch_bp<A, Bv>(im, table, dviews, nview, g, d, stream); continue; }
        CTP_BP_CASE(1, 1) CTP_BP_CASE(1, 2) CTP_BP_CASE(2, 1) CTP_BP_CASE(2, 2)
        CTP_BP_CASE(2, 3) CTP_BP_CASE(3, 2) CTP_BP_CASE(3, 3) CTP_BP_CASE(1, 3) CTP_BP_CASE(3, 1)
        CTP_BP_CASE(1, 4) CTP_BP_CASE(4, 1) CTP_BP_CASE(2, 4) CTP_BP_CASE(4, 2) CTP_BP_CASE(3, 4)
        CTP_BP_CASE(4, 3) CTP_BP_CASE(4, 4)
#undef CTP_BP_CASE
        fail("cone backprojection: unsupported window");
    }
    // host plan vector must outlive the async copy
    CTP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// Volumes of 2^29 voxels or more (up to the 2^31 - 1 the reference's int
// indices allow, cuda/siddon.cu:166-167): the kernels address the volume with
// 32-bit byte offsets, so such volumes are projected as z-slabs of whole
// slices, each a grid of its own with its centre moved to the slab's centre.
// Line integrals and their transposes add over slabs: prj += A_s img_s for
// every slab s, img_s += A_s^T prj.  Slab boundaries are voxel planes of the
// full grid, so the chords are the full grid's up to FP32 rounding.
// ---------------------------------------------------------------------------
static int cone_slab_slices(const Grid& g) {
    const size_t per = (size_t)g.nx * g.ny;
    const size_t lim = (1ull << 29) - 1;
    int ns = g.nz;
    if ((size_t)g.nz * per > lim || (size_t)(g.nz + 3) * per > lim) {
        if (per * 4 > lim) fail("cone projection: a slice of more than 2^27 voxels is not supported");
        ns = (int)(lim / per) - 3;
    }
    // cone_slab_slices: at most this many slices per slab (0 = only where needed)
    const int t = tuning_get("cone_slab_slices", 0);
    if (t < 0) fail("cone_slab_slices must be positive");
    return t > 0 ? std::min(ns, t) : ns;
}

static Grid cone_slab(const Grid& g, int z0, int ns) {
    Grid s = g;
    s.nz = ns;
    s.cz = (float)((double)g.cz + ((double)z0 + ns / 2.0 - g.nz / 2.0) * (double)g.dz);
    return s;
}

void siddon_cone_fp(float* prj, const float* img, const ConeGeomDev& geo, int B, const Grid& g, const Det& d,
                    int nview, hipStream_t stream) {
    const int ns = cone_slab_slices(g);
    if (ns >= g.nz) return siddon_cone_fp_grid(prj, img, geo, B, g, d, nview, stream);
    const size_t per = (size_t)g.nx * g.ny, nimg = per * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    for (int b = 0; b < B; b++)
        for (int z0 = 0; z0 < g.nz; z0 += ns) {
            const int n = std::min(ns, g.nz - z0);
            siddon_cone_fp_grid(prj + b * nprj, img + b * nimg + (size_t)z0 * per, geo, 1, cone_slab(g, z0, n), d,
                                nview, stream);
        }
}

void siddon_cone_bp(float* img, const float* prj, const ConeGeomDev& geo, const ConeGeomHost& hgeo, int B,
                    const Grid& g, const Det& d, int nview, hipStream_t stream) {
    const int ns = cone_slab_slices(g);
    if (ns >= g.nz) return siddon_cone_bp_grid(img, prj, geo, hgeo, B, g, d, nview, stream);
    const size_t per = (size_t)g.nx * g.ny, nimg = per * g.nz, nprj = (size_t)d.nu * d.nv * nview;
    for (int b = 0; b < B; b++)
        for (int z0 = 0; z0 < g.nz; z0 += ns) {
            const int n = std::min(ns, g.nz - z0);
            siddon_cone_bp_grid(img + b * nimg + (size_t)z0 * per, prj + b * nprj, geo, hgeo, 1, cone_slab(g, z0, n),
                                d, nview, stream);
        }
    CTP_CHECK(hipStreamSynchronize(stream));
}

// ===========================================================================
// Batched plans (no counterpart in the reference): many cone-beam jobs --
// volumes of different sizes, each with its own trajectory and detector -- in
// one persistent launch per direction.  Everything that depends only on the
// geometry (device copies of the geometry, swap flags, FP64 projective maps,
// per-ray length x hit factors, column tables, the rare near-parallel rays) is
// prepared once when the plan is created; forward() and backward() then only
// launch kernels on the caller's stream: no allocation, host synchronisation,
// copy or memset, so they can be captured into a HIP graph and replayed.
// ===========================================================================
struct ConeJobDev {
    Grid g;
    Det d;
    int nview;
    float* img;               // bound volume [nz][ny][nx]
    float* prj;               // bound projections [nview][nv][nu]
    float* imgT;              // transposed copy for x-dominant views (null if none)
    float* imgz;              // z-contiguous padded copy for the column-driven form
    int zcol;                 // floats per z column of imgz (nz + 3)
    const float *dc, *du, *dv, *src;
    const int* swapf;
    const ConeViewBP* views;
    float* wtab;              // per call: prj * lhit, window layout of height wv
    const float* lhit;        // per ray: |q - S| if the ray hits the grid (and is regular), else 0
    const float2* utab;
    const float2* vtab;
    int nub, ntile;           // forward tiles per view
    int nbx, nby, nbz;        // backprojection blocks
    int wu, wv;               // backprojection candidate window of this job
    long fp_begin, bp_begin;  // first block of the job in the batched launches
};

// Per-ray geometry of the column mode (FP64): ray length, hit, near-parallel
// exception, 1/w_x, 1/w_y (column) and 1/w_z (row).  The same rules as
// siddon_bp_coltable_kernel.
__device__ inline void col_ray(long i, const float* __restrict__ pdc, const float* __restrict__ pdu,
                               const float* __restrict__ pdv, const float* __restrict__ psrc, const Grid& g,
                               const Det& d, double& L, bool& hit, bool& exception, float2& uinv, float2& vinv) {
    const int iu = (int)(i % d.nu);
    const int iv = (int)((i / d.nu) % d.nv);
    const int view = (int)(i / ((long)d.nu * d.nv));
    const double u = ((double)iu - d.off_u - (d.nu - 1) / 2.0) * d.du;
    const double v = ((double)iv - d.off_v - (d.nv - 1) / 2.0) * d.dv;
    double S[3], Q[3];
    for (int k = 0; k < 3; k++) {
        S[k] = psrc[3 * view + k];
        Q[k] = pdc[3 * view + k] + pdu[3 * view + k] * u + pdv[3 * view + k] * v;
    }
    const double dd[3] = {g.dx, g.dy, g.dz};
    const double cc[3] = {g.cx, g.cy, g.cz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    double Sm[3] = {S[0], S[1], S[2]}, Qm[3] = {Q[0], Q[1], Q[2]};
    {
        double ex = g.nx * (double)g.dx, ey = g.ny * (double)g.dy, ez = g.nz * (double)g.dz;
        double r = sqrt(ex * ex + ey * ey + ez * ez);
        double w[3] = {Q[0] - S[0], Q[1] - S[1], Q[2] - S[2]};
        double len = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(len < 1e-6)) for (int k = 0; k < 3; k++) w[k] /= len;
        double t0 = 0, t1 = 0;
        for (int k = 0; k < 3; k++) { t0 += (cc[k] - S[k]) * w[k]; t1 += (Q[k] - cc[k]) * w[k]; }
        if (t0 > r) for (int k = 0; k < 3; k++) Sm[k] = S[k] + (t0 - r) * w[k];
        if (t1 > r) for (int k = 0; k < 3; k++) Qm[k] = Q[k] - (t1 - r) * w[k];
    }
    double wi[3], di[3], smi[3];
    bool rpar[3], cpar[3];
    L = 0;
    for (int k = 0; k < 3; k++) {
        L += (Q[k] - S[k]) * (Q[k] - S[k]);
        wi[k] = (Q[k] - S[k]) / dd[k];
        di[k] = (Qm[k] - Sm[k]) / dd[k];
        smi[k] = (Sm[k] - cc[k]) / dd[k] + nn[k] / 2.0;
        rpar[k] = !(fabs(di[k]) > 1e-6);
        cpar[k] = (Q[k] - S[k]) == 0.0;
    }
    L = sqrt(L);
    double amin = 0, amax = 1;
    hit = true;
    for (int k = 0; k < 3; k++) {
        if (!rpar[k]) {
            double a0 = (0 - smi[k]) / di[k], a1 = (nn[k] - smi[k]) / di[k];
            amin = fmax(amin, fmin(a0, a1));
            amax = fmin(amax, fmax(a0, a1));
        } else if (smi[k] < 0 || smi[k] > nn[k]) {
            hit = false;
        }
    }
    if (!(amax > amin)) hit = false;
    for (int k = 0; k < 3; k++)
        if (rpar[k]) {
            int fixed = (int)(smi[k] + amin * di[k]);
            if (fixed < 0 || fixed >= nn[k]) hit = false;
        }
    uinv = make_float2(cpar[0] ? 0.f : (float)(1.0 / wi[0]), cpar[1] ? 0.f : (float)(1.0 / wi[1]));
    const float qz = cpar[2] ? 0.f : (float)(1.0 / wi[2]);
    vinv = make_float2(qz, 0.5f * fabsf(qz));
    exception = rpar[0] != cpar[0] || rpar[1] != cpar[1] || rpar[2] != cpar[2];
}

// plan creation: lhit, utab, vtab and the list of exception rays (count only
// when list is null)
__global__ __launch_bounds__(256) void siddon_plan_rays_kernel(float* __restrict__ lhit, float2* __restrict__ utab,
                                                               float2* __restrict__ vtab,
                                                               ConeViewBP* __restrict__ views, int* __restrict__ nexc,
                                                               int* __restrict__ exc, const float* __restrict__ pdc,
                                                               const float* __restrict__ pdu,
                                                               const float* __restrict__ pdv,
                                                               const float* __restrict__ psrc, Grid g, Det d, int nview) {
    const long n = (long)d.nu * d.nv * nview;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double L;
    bool hit, ex;
    float2 ui, vi;
    col_ray(i, pdc, pdu, pdv, psrc, g, d, L, hit, ex, ui, vi);
    const int iu = (int)(i % d.nu), iv = (int)((i / d.nu) % d.nv), view = (int)(i / ((long)d.nu * d.nv));
    if (iv == 0) utab[(long)view * d.nu + iu] = ui;
    if (iu == 0) vtab[(long)view * d.nv + iv] = vi;
    // zero entries (axis-parallel column or row): parallel-ray handling in this view
    if ((iv == 0 && (ui.x == 0.f || ui.y == 0.f)) || (iu == 0 && vi.x == 0.f)) views[view].flags = 1;
    lhit[i] = (hit && !ex) ? (float)L : 0.f;
    if (ex) {
        const int k = atomicAdd(nexc, 1);
        if (exc) exc[k] = (int)i;
    }
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// weights prj * lhit in the window layout of the column backprojector
// grid: (column blocks, rows, views)
__global__ __launch_bounds__(256) void window_weights_kernel(float* __restrict__ w, const float* __restrict__ prj,
                                                             const float* __restrict__ lhit, int nu, int nv,
                                                             int dup) {
    const int iu = blockIdx.x * 256 + threadIdx.x, iv = blockIdx.y;
    if (iu >= nu) return;
    const size_t i = ((size_t)blockIdx.z * nv + iv) * nu + iu;
    float* out = w + i * dup;
    for (int jv = 0; jv < dup; jv++) {
        const size_t r = i + (size_t)jv * nu;
        out[jv] = iv + jv < nv ? prj[r] * lhit[r] : 0.f;
    }
}

// ray-driven scatter of the listed rays (the forward projector's traversal),
// img += p * len
__global__ __launch_bounds__(256) void siddon_scatter_rays_kernel(float* __restrict__ img, const float* __restrict__ prj,
                                                                  const int* __restrict__ list, int n,
                                                                  const float* __restrict__ pdc,
                                                                  const float* __restrict__ pdu,
                                                                  const float* __restrict__ pdv,
                                                                  const float* __restrict__ psrc, Grid g, Det d) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int i = list[k];
    const int iu = i % d.nu, iv = (i / d.nu) % d.nv, view = i / (d.nu * d.nv);
    const float p = prj[i];
    if (p == 0.f) return;
    SiddonRay r;
    if (!siddon_setup(r, ld3(psrc, view), cell_centre(pdc, pdu, pdv, view, iu, iv, d), g, false)) return;
    bool inside = true;
    while (inside) {
        const int off = r.off;
        const float len = siddon_step(r, inside);
        atomicAdd(img + (off >> 2), p * len);
    }
}

__global__ void reset_counters_kernel(int* __restrict__ ctr, int n) {
    if ((int)threadIdx.x < n) ctr[threadIdx.x] = 0;
}

template <int TU, int TV>
__global__ __launch_bounds__(256) void siddon_cone_fp_batch_kernel(const ConeJobDev* __restrict__ jobs, int njobs,
                                                                   long total, int xcd_order, int* __restrict__ ctr,
                                                                   int persist) {
    __shared__ int slot;
    for (int b = first_block(ctr, 8, &slot, persist); b < total; b = following_block(ctr, 8, &slot, persist)) {
        int j = 0;
        while (j + 1 < njobs && b >= jobs[j + 1].fp_begin) j++;
        const ConeJobDev* J = jobs + j;
        int view, tile;
        fp_tile_of(b - J->fp_begin, J->ntile, xcd_order, view, tile);
        if (tile >= J->ntile || view >= J->nview) continue;
        siddon_fp_block<TU, TV, false>(J->prj, J->img, J->imgT, J->swapf[view], J->dc, J->du, J->dv, J->src, J->g,
                                       J->d, view, tile % J->nub, tile / J->nub);
    }
}

// column-driven form: a block is 4 adjacent detector columns of one view.
// Five waves per SIMD as the single-volume kernel gets (the job lookup would
// otherwise push it to 100 VGPRs and four waves: 17% slower).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void siddon_cone_fp_batch_col_kernel(const ConeJobDev* __restrict__ jobs, int njobs,
                                                                       long total, int xcd_order,
                                                                       int* __restrict__ ctr, int persist) {
    __shared__ int slot;
    for (int b = first_block(ctr, 8, &slot, persist); b < total; b = following_block(ctr, 8, &slot, persist)) {
        int j = 0;
        while (j + 1 < njobs && b >= jobs[j + 1].fp_begin) j++;
        const ConeJobDev* J = jobs + j;
        int view, grp;
        fp_tile_of(b - J->fp_begin, J->ntile, xcd_order, view, grp);
        const int iu = grp * 4 + (threadIdx.x >> 6);
        if (grp >= J->ntile || view >= J->nview || iu >= J->d.nu) continue;
        const Grid g = J->g;
        const __amdgpu_buffer_rsrc_t volz =
            __builtin_amdgcn_make_buffer_rsrc((void*)J->imgz, (short)0, 4 * J->zcol * g.nx * g.ny, 0x00020000);
        fp_column<false, 1>(J->prj, volz, J->img, J->dc, J->du, J->dv, J->src, g, J->d, view, iu, J->zcol, 0, 0);
    }
}

// Eight waves per SIMD for z-runs up to 4, as the single-volume kernel gets
// (the job lookup would otherwise cost registers and occupancy).
template <int WU, int WV, int NZR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NZR <= 4 ? 8 : 1))) void siddon_cone_bp_batch_kernel(const ConeJobDev* __restrict__ jobs, int njobs,
                                                                   long total, int* __restrict__ ctr, int persist) {
    __shared__ int slot;
    for (int b = first_block(ctr, 1, &slot, persist); b < total; b = following_block(ctr, 1, &slot, persist)) {
        int j = 0;
        while (j + 1 < njobs && b >= jobs[j + 1].bp_begin) j++;
        const ConeJobDev* J = jobs + j;
        const long local = b - J->bp_begin;
        const int bx = (int)(local % J->nbx);
        const long rest = local / J->nbx;
        const int by = (int)(rest % J->nby), bz = (int)(rest / J->nby);
        const WTabs wt{{J->wtab, nullptr, nullptr, nullptr}};
        // the group's kernel is compiled for its exact window (WU, WV) = (wu, wv);
        // its loads stay voxel by voxel, as in siddon_cone_bp_plan_kernel
        bp_col_block<WU, WV, NZR, 1, false>(J->img, 0, wt, J->utab, J->vtab, J->views, J->nview, J->g, J->d.nu,
                                            J->d.nv, bx, by, bz);
    }
}

// The whole backward of a plan in one persistent launch: blocks are numbered
// over all jobs (job j's from jobs[j].bp_begin) and each runs the column-mode
// backprojector compiled for its job's exact candidate window (a switch on
// the job's window, uniform across the wavefront), so no job does the work of
// a larger window; the registers are those of the largest window's code
// (107 VGPRs at a z-run of 8, as the per-shape kernels of the 2x2 and 3x3
// windows).  Measured over four volumes (128^3 ... 256^3, 720 views): 47.6
// against 48.1 ms for one launch per shape at a z-run of 8, 49.3 against 50.9
// at 4.  Loads stay voxel by voxel: issued ahead, the backward measured 1%
// slower (48.96 against 48.45 ms with one launch per shape; 47.65 against
// 47.45 ms here with the windows up to 3x3 issuing them ahead).
template <int NZR>
__global__ __launch_bounds__(256) void siddon_cone_bp_plan_kernel(const ConeJobDev* __restrict__ jobs, int njobs,
                                                                  long total, int* __restrict__ ctr, int persist) {
    __shared__ int slot;
    for (int b = first_block(ctr, 1, &slot, persist); b < total; b = following_block(ctr, 1, &slot, persist)) {
        int j = 0;
        while (j + 1 < njobs && b >= jobs[j + 1].bp_begin) j++;
        const ConeJobDev* J = jobs + j;
        const long local = b - J->bp_begin;
        const int bx = (int)(local % J->nbx);
        const long rest = local / J->nbx;
        const int by = (int)(rest % J->nby), bz = (int)(rest / J->nby);
        const WTabs wt{{J->wtab, nullptr, nullptr, nullptr}};
        switch (J->wu * 8 + J->wv) {
#define CTP_BPP(A, Bv)                                                                                          \
    case A * 8 + Bv:                                                                                            \
        bp_col_block<A, Bv, NZR, 1, false>(J->img, 0, wt, J->utab, J->vtab, J->views, J->nview, J->g, J->d.nu, \
                                           J->d.nv, bx, by, bz);                                                \
        break;
            CTP_BPP(1, 1) CTP_BPP(1, 2) CTP_BPP(2, 1) CTP_BPP(2, 2) CTP_BPP(1, 3) CTP_BPP(3, 1) CTP_BPP(2, 3)
            CTP_BPP(3, 2) CTP_BPP(3, 3) CTP_BPP(1, 4) CTP_BPP(4, 1) CTP_BPP(2, 4) CTP_BPP(4, 2) CTP_BPP(3, 4)
            CTP_BPP(4, 3) CTP_BPP(4, 4)
#undef CTP_BPP
            default: break;
        }
    }
}

template <typename K>
static int persistent_grid(K kernel, long total) {
    int dev = 0, nb = 0;
    CTP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CTP_CHECK(hipGetDeviceProperties(&prop, dev));
    CTP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kernel), 256, 0));
    long grid = (long)std::max(nb, 1) * prop.multiProcessorCount;
    grid = std::max(8L, grid / 8 * 8);                   // a multiple of 8: XCD of a block is fixed
    return (int)std::min(grid, (total + 7) / 8 * 8);
}

// 1: persistent blocks with work counters, 2: one hardware-dispatched block per
// tile (cone_plan_persistent)
static int plan_persist() { return tuning_get("cone_plan_persistent", 1) == 1 ? 1 : 0; }

template <typename K>
static int plan_grid(K kernel, long total) {
    return plan_persist() ? persistent_grid(kernel, total) : (int)total;
}

struct ConePlan::Impl {
    std::vector<ConeJobSpec> spec;
    std::vector<ConeJobDev> jobs;
    std::vector<void*> mem;
    std::vector<int> nexc;
    std::vector<int*> exc;
    std::vector<bool> need_t;
    ConeJobDev* djobs = nullptr;
    // backprojection by window shape (the per-shape launches of
    // cone_plan_bp_launch 2, blocks numbered within the group); the default
    // single launch numbers blocks over all jobs (ConeJobDev::bp_begin of djobs)
    struct BPGroup {
        int wu, wv;
        ConeJobDev* djobs;
        int njobs;
        long total;
    };
    std::vector<BPGroup> groups;
    int* ctr = nullptr;           // work counters of the persistent launches (8 forward, one per backward group)
    int tu = 0, order = 1, nzr = 4;   // tu 0: column-driven forward form
    long fp_total = 0, bp_total = 0;   // bp_total: blocks of the backward over all jobs

    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        CTP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        mem.push_back(p);
        return static_cast<T*>(p);
    }
    ~Impl() {
        for (void* p : mem) (void)hipFree(p);
    }
    void layout();                 // block offsets for the current tuning; uploads the job table
    void forward(hipStream_t s);
    void backward(hipStream_t s);
};

void ConePlan::Impl::layout() {
    // block indices (and the counters handing them out, which run past the end
    // by one grid) must fit in 31 bits
    const long limit = 0x7fffffffL - (1L << 24);
    fp_total = bp_total = 0;
    long bp_all = 0;
    for (auto& J : jobs) {
        if (tu == 0) {                      // groups of 4 detector columns
            J.nub = 0;
            J.ntile = ceil_div(J.d.nu, 4);
        } else {
            const int TV = 64 / tu;
            J.nub = ceil_div(J.d.nu, tu);
            J.ntile = J.nub * ceil_div(J.d.nv, 4 * TV);
        }
        J.nbx = ceil_div(J.g.nx, 16);
        J.nby = ceil_div(J.g.ny, 16);
        J.nbz = ceil_div(J.g.nz, nzr);
        J.fp_begin = fp_total;
        J.bp_begin = bp_all;
        fp_total += (long)J.nview * 8 * ((J.ntile + 7) / 8);
        bp_all += (long)J.nbx * J.nby * J.nbz;
    }
    if (fp_total > limit || bp_all > limit) fail("cone plan: too many blocks for one batch; split the jobs");
    CTP_CHECK(hipMemcpy(djobs, jobs.data(), sizeof(ConeJobDev) * jobs.size(), hipMemcpyHostToDevice));
    bp_total = 0;
    for (BPGroup& G : groups) {
        std::vector<ConeJobDev> gj;
        G.total = 0;
        for (const ConeJobDev& J : jobs) {
            if (J.wu != G.wu || J.wv != G.wv) continue;
            ConeJobDev c = J;
            c.bp_begin = G.total;
            G.total += (long)J.nbx * J.nby * J.nbz;
            gj.push_back(c);
        }
        if (G.total > limit) fail("cone plan: too many blocks for one batch; split the jobs");
        bp_total += G.total;
        CTP_CHECK(hipMemcpy(G.djobs, gj.data(), sizeof(ConeJobDev) * gj.size(), hipMemcpyHostToDevice));
    }
}

void ConePlan::Impl::forward(hipStream_t s) {
    for (size_t j = 0; j < jobs.size(); j++) {
        const ConeJobDev& J = jobs[j];
        if (tu == 0) {
            transpose_zcol_kernel<<<dim3(ceil_div(J.g.nx, 32), ceil_div(J.g.nz, 32), J.g.ny), 256, 0, s>>>(
                J.imgz, J.img, J.g.nx, J.g.ny, J.g.nz, J.zcol);
            CTP_LAUNCH_CHECK();
        } else if (need_t[j]) {
            transpose_xy_kernel<<<dim3(ceil_div(J.g.nx, 32), ceil_div(J.g.ny, 32), J.g.nz), 256, 0, s>>>(
                J.imgT, J.img, J.g.nx, J.g.ny);
            CTP_LAUNCH_CHECK();
        }
    }
    const int nj = (int)jobs.size();
    reset_counters_kernel<<<1, 64, 0, s>>>(ctr, 8);
    CTP_LAUNCH_CHECK();
    TimedLaunch t("siddon_cone_fp_batch", s);
    if (tu == 0) {
        auto k = siddon_cone_fp_batch_col_kernel;
        k<<<plan_grid(k, fp_total), 256, 0, s>>>(djobs, nj, fp_total, order, ctr, plan_persist());
        CTP_LAUNCH_CHECK();
        return;
    }
#define CTP_FPB(A, Bv)                                                                                     \
    if (tu == A) {                                                                                        \
        auto k = siddon_cone_fp_batch_kernel<A, Bv>;                                                      \
        k<<<plan_grid(k, fp_total), 256, 0, s>>>(djobs, nj, fp_total, order, ctr, plan_persist());        \
    }
    CTP_FPB(8, 8) CTP_FPB(16, 4) CTP_FPB(32, 2) CTP_FPB(64, 1)
#undef CTP_FPB
    CTP_LAUNCH_CHECK();
}

void ConePlan::Impl::backward(hipStream_t s) {
    for (size_t j = 0; j < jobs.size(); j++) {
        const ConeJobDev& J = jobs[j];
        const size_t nimg = (size_t)J.g.nx * J.g.ny * J.g.nz;
        fill_zero_kernel<<<(unsigned)ceil_div((long)nimg, 256), 256, 0, s>>>(J.img, nimg);
        CTP_LAUNCH_CHECK();
        window_weights_kernel<<<dim3(ceil_div(J.d.nu, 256), J.d.nv, J.nview), 256, 0, s>>>(J.wtab, J.prj, J.lhit,
                                                                                            J.d.nu, J.d.nv, J.wv);
        CTP_LAUNCH_CHECK();
        if (nexc[j] > 0) {
            siddon_scatter_rays_kernel<<<ceil_div(nexc[j], 256), 256, 0, s>>>(J.img, J.prj, exc[j], nexc[j], J.dc, J.du,
                                                                              J.dv, J.src, J.g, J.d);
            CTP_LAUNCH_CHECK();
        }
    }
    reset_counters_kernel<<<1, 64, 0, s>>>(ctr + 8, (int)groups.size());
    CTP_LAUNCH_CHECK();
    TimedLaunch t("siddon_cone_bp_batch", s);
    if (tuning_get("cone_plan_bp_launch", 1) == 1) {
        const int nj = (int)jobs.size();
#define CTP_BPP_LAUNCH(Z)                                                                               \
    if (nzr == Z) {                                                                                     \
        auto k = siddon_cone_bp_plan_kernel<Z>;                                                         \
        k<<<plan_grid(k, bp_total), 256, 0, s>>>(djobs, nj, bp_total, ctr + 8, plan_persist());         \
    }
        CTP_BPP_LAUNCH(2) CTP_BPP_LAUNCH(4) CTP_BPP_LAUNCH(8)
#undef CTP_BPP_LAUNCH
        CTP_LAUNCH_CHECK();
        return;
    }
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const BPGroup& G = groups[gi];
        int* gc = ctr + 8 + gi;
#define CTP_BPB(A, Bv, Z)                                                                                  \
    if (G.wu == A && G.wv == Bv && nzr == Z) {                                                            \
        auto k = siddon_cone_bp_batch_kernel<A, Bv, Z>;                                                   \
        k<<<plan_grid(k, G.total), 256, 0, s>>>(G.djobs, G.njobs, G.total, gc, plan_persist());           \
    }
#define CTP_BPB_Z(A, Bv) CTP_BPB(A, Bv, 2) CTP_BPB(A, Bv, 4) CTP_BPB(A, Bv, 8)
        CTP_BPB_Z(1, 1) CTP_BPB_Z(1, 2) CTP_BPB_Z(2, 1) CTP_BPB_Z(2, 2) CTP_BPB_Z(1, 3) CTP_BPB_Z(3, 1)
        CTP_BPB_Z(2, 3) CTP_BPB_Z(3, 2) CTP_BPB_Z(3, 3) CTP_BPB_Z(1, 4) CTP_BPB_Z(4, 1) CTP_BPB_Z(2, 4)
        CTP_BPB_Z(4, 2) CTP_BPB_Z(3, 4) CTP_BPB_Z(4, 3) CTP_BPB_Z(4, 4)
#undef CTP_BPB_Z
#undef CTP_BPB
        CTP_LAUNCH_CHECK();
    }
}

ConePlan::ConePlan(const std::vector<ConeJobSpec>& specs, bool autotune) : impl_(new Impl) {
    Impl& P = *impl_;
    P.spec = specs;
    for (const ConeJobSpec& S : specs) {
        const Grid& g = S.g;
        const Det& d = S.d;
        const int nview = S.nview;
        const size_t nimg = (size_t)g.nx * g.ny * g.nz, nprj = (size_t)d.nu * d.nv * nview;
        if (nimg >= (1ull << 29) || (size_t)g.nx * g.ny * (g.nz + 3) >= (1ull << 29))
            fail("cone plan: volumes of 2^29 voxels or more are not supported");
        ConeGeomHost hg{S.det_center, S.det_u, S.det_v, S.src};
        ConeBPPlan bp = make_cone_bp_plan(hg, g, d, nview);
        if (!column_mode(hg, bp, nview))
            fail("cone plan: every view must have det_v along z and det_u horizontal (column geometry)");
        if (d.nu >= (1 << 20)) fail("cone plan: detectors of 2^20 columns or more are not supported");
        if (d.nv > 65535 || nview > 65535) fail("cone plan: more than 65535 detector rows or views");
        if (bp.wu > 4 || bp.wv > 4)
            fail("cone plan: voxel shadow spans more than 4 detector cells; use finer voxels or coarser cells");
        ConeJobDev J{};
        J.wu = std::max(1, std::min(bp.wu, d.nu));
        J.wv = std::max(1, std::min(bp.wv, d.nv));
        J.g = g;
        J.d = d;
        J.nview = nview;
        J.img = S.img;
        J.prj = S.prj;
        float* geo = P.alloc<float>(12 * (size_t)nview);
        const float* hs[4] = {S.det_center, S.det_u, S.det_v, S.src};
        for (int k = 0; k < 4; k++)
            CTP_CHECK(hipMemcpy(geo + 3 * k * (size_t)nview, hs[k], 12 * (size_t)nview, hipMemcpyHostToDevice));
        J.dc = geo; J.du = geo + 3 * nview; J.dv = geo + 6 * nview; J.src = geo + 9 * nview;
        // swap flags from the host geometry (as siddon_swap_flags_kernel)
        std::vector<int> sw(nview);
        bool any = false;
        for (int i = 0; i < nview; i++) {
            const float wx = S.det_center[3 * i] - S.src[3 * i], wy = S.det_center[3 * i + 1] - S.src[3 * i + 1];
            sw[i] = std::fabs(wx) > std::fabs(wy) ? 1 : 0;
            any |= sw[i] != 0;
        }
        int* dsw = P.alloc<int>(nview);
        CTP_CHECK(hipMemcpy(dsw, sw.data(), sizeof(int) * nview, hipMemcpyHostToDevice));
        J.swapf = dsw;
        J.imgT = any ? P.alloc<float>(nimg) : nullptr;
        P.need_t.push_back(any);
        J.zcol = g.nz + 3;
        J.imgz = P.alloc<float>((size_t)g.nx * g.ny * J.zcol);
        ConeViewBP* dv = P.alloc<ConeViewBP>(nview);
        for (ConeViewBP& cv : bp.views) cv.flags = 0;        // set by siddon_plan_rays_kernel
        CTP_CHECK(hipMemcpy(dv, bp.views.data(), sizeof(ConeViewBP) * nview, hipMemcpyHostToDevice));
        J.views = dv;
        float* lhit = P.alloc<float>(nprj);
        float2* ut = P.alloc<float2>((size_t)nview * d.nu);
        float2* vt = P.alloc<float2>((size_t)nview * d.nv + 4);   // padded: row-table loads
        int* cnt = P.alloc<int>(1);
        CTP_CHECK(hipMemset(cnt, 0, sizeof(int)));
        siddon_plan_rays_kernel<<<ceil_div((long)nprj, 256), 256>>>(lhit, ut, vt, dv, cnt, nullptr, J.dc, J.du, J.dv, J.src,
                                                                   g, d, nview);
        CTP_LAUNCH_CHECK();
        int n = 0;
        CTP_CHECK(hipMemcpy(&n, cnt, sizeof(int), hipMemcpyDeviceToHost));
        int* list = nullptr;
        if (n > 0) {
            list = P.alloc<int>(n);
            CTP_CHECK(hipMemset(cnt, 0, sizeof(int)));
            siddon_plan_rays_kernel<<<ceil_div((long)nprj, 256), 256>>>(lhit, ut, vt, dv, cnt, list, J.dc, J.du, J.dv,
                                                                       J.src, g, d, nview);
            CTP_LAUNCH_CHECK();
        }
        P.nexc.push_back(n);
        P.exc.push_back(list);
        J.lhit = lhit;
        J.utab = ut;
        J.vtab = vt;
        P.jobs.push_back(J);
    }

    // weight tables in the window layout of each job's window height
    for (ConeJobDev& J : P.jobs) J.wtab = P.alloc<float>((size_t)J.d.nu * J.d.nv * J.nview * J.wv + 16);
    P.djobs = P.alloc<ConeJobDev>(P.jobs.size());
    for (const ConeJobDev& J : P.jobs) {
        bool found = false;
        for (Impl::BPGroup& G : P.groups) found |= G.wu == J.wu && G.wv == J.wv;
        if (found) continue;
        int n = 0;
        for (const ConeJobDev& K : P.jobs) n += K.wu == J.wu && K.wv == J.wv;
        P.groups.push_back(Impl::BPGroup{J.wu, J.wv, P.alloc<ConeJobDev>(n), n, 0});
    }
    P.ctr = P.alloc<int>(8 + P.groups.size());
    const int mode = tuning_get("cone_fp_mode", 0);
    if (mode != 0 && mode != 1) fail("cone_fp_mode must be 0 or 1");
    P.tu = mode == 0 ? 0 : tuning_get("cone_fp_tile_u", 32);
    P.order = fp_xcd_order(mode);
    P.nzr = tuning_get("cone_bp_zrun", 4);
    if (P.tu != 0 && P.tu != 8 && P.tu != 16 && P.tu != 32 && P.tu != 64)
        fail("cone_fp_tile_u must be 8, 16, 32 or 64");
    if (P.nzr != 2 && P.nzr != 4 && P.nzr != 8) fail("cone plan: cone_bp_zrun must be 2, 4 or 8");
    P.layout();
    CTP_CHECK(hipDeviceSynchronize());
    if (autotune) tune();
}

ConePlan::~ConePlan() = default;

void ConePlan::forward(hipStream_t s) { impl_->forward(s); }
void ConePlan::backward(hipStream_t s) { impl_->backward(s); }
void ConePlan::config(int& tile_u, int& xcd_order, int& zrun) const {
    tile_u = impl_->tu;
    xcd_order = impl_->order;
    zrun = impl_->nzr;
}

// Autotuner: times every forward form (column driven, or ray per lane with
// each tile shape) with each block order, and every backprojection z-run, on
// the plan's own buffers (one warm-up and two timed runs each, HIP events on
// the null stream) and keeps the fastest.  The bound buffers' contents are
// overwritten.
void ConePlan::tune() {
    Impl& P = *impl_;
    struct OwnEvents {   // released on every exit
        hipEvent_t a = nullptr, b = nullptr;
        ~OwnEvents() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    CTP_CHECK(hipEventCreate(&ev.a));
    CTP_CHECK(hipEventCreate(&ev.b));
    hipEvent_t a = ev.a, b = ev.b;
    auto time = [&](auto&& fn) {
        fn();
        CTP_CHECK(hipEventRecord(a, nullptr));
        fn();
        fn();
        CTP_CHECK(hipEventRecord(b, nullptr));
        CTP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        CTP_CHECK(hipEventElapsedTime(&ms, a, b));
        return ms;
    };
    float best = 1e30f;
    int btu = P.tu, bor = P.order;
    for (int tu : {0, 16, 32, 64})
        for (int order : {1, 0}) {
            P.tu = tu;
            P.order = order;
            P.layout();
            const float ms = time([&] { P.forward(nullptr); });
            if (ms < best) { best = ms; btu = tu; bor = order; }
        }
    P.tu = btu;
    P.order = bor;
    best = 1e30f;
    int bz = P.nzr;
    for (int z : {2, 4, 8}) {
        P.nzr = z;
        P.layout();
        const float ms = time([&] { P.backward(nullptr); });
        if (ms < best) { best = ms; bz = z; }
    }
    P.nzr = bz;
    P.layout();
    CTP_CHECK(hipDeviceSynchronize());
}

}  // namespace ctp
