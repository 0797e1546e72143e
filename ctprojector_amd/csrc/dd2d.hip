// Distance-driven fan-beam (equiangular) and parallel-beam projectors on
// MI355X (gfx950): cuda/distanceDriven2DBoxInt.cu, cuda/distanceDriven2DBranchless.cu,
// cuda/distanceDriven2DUtils.cu of the reference.
//
// The reference projector is a stack of 2D problems: a ray's footprint on an
// image row depends on (view, detector column) only, and a pixel's footprint on
// the detector on (view, x, y) only; z enters through a per-detector-row linear
// interpolation (FP) or a per-slice row lookup (BP).  Here every image slice of
// every batch item is a "channel":
//
//   FP: one lane per (view, column) carries CH+1 channel accumulators; for each
//       image row it computes the footprint weights once and applies them to a
//       contiguous channel vector of the image stored channel-last
//       ([y][x][C] for y-main views, [x][y][C] for x-main views).  The detector
//       rows are produced at the end from the accumulators (z interpolation).
//   BP: one lane per pixel (x, y) carries CH channel accumulators; for each view
//       it computes the detector footprint once and applies it to a contiguous
//       channel vector of the pre-weighted projections stored
//       [view][u][C] (one channel per image slice, the slice's detector row
//       already selected).  No atomics: every pixel column has one owner.
//
// Decisions that are discontinuous in the view angle (the 45-degree main-axis
// switch of the backprojector and of the BP pre-weight) are taken in FP64 on
// the host so that they agree with the FP64 oracle.
#include "common.h"
#include "runtime.h"
#include "ops.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace ctp {

struct DDView {
    float c, s;        // cos / sin of the view angle (FP64 rounded to FP32)
    float deg;         // the view angle
    float wx;          // blend weight of the y-main trace, clamp((|c|-|s|+m)/2m, 0, 1)
    float rev;         // parallel FP with force-FBP: 1/(dy/|c|) or 1/(dx/|s|)
    int flags;         // bit0: FP y-main trace, bit1: FP x-main trace, bit2: BP y-main (|c| > |s|)
    float pad0, pad1;
};

static std::vector<DDView> make_dd_views(const float* deg, int nview, const Grid& g) {
    std::vector<DDView> v(nview);
    const double m = (double)kDDMargin;
    for (int i = 0; i < nview; i++) {
        const double d = deg[i];
        const double c = std::cos(d), s = std::sin(d);
        double wx = (std::fabs(c) - std::fabs(s) + m) / (2 * m);
        wx = std::min(1.0, std::max(0.0, wx));
        DDView& o = v[i];
        o.c = (float)c;
        o.s = (float)s;
        o.deg = deg[i];
        o.wx = (float)wx;
        int f = 0;
        if (std::fabs(c) > std::fabs(s) - m) f |= 1;
        if (std::fabs(c) < std::fabs(s) + m) f |= 2;
        const bool ymain = std::fabs(c) > std::fabs(s);
        if (ymain) f |= 4;
        o.rev = ymain ? (float)(std::fabs(c) / g.dy) : (float)(std::fabs(s) / g.dx);
        o.flags = f;
        o.pad0 = o.pad1 = 0;
    }
    return v;
}

// The branchless projectors switch main axes at 45 degrees without a blend
// (cuda/distanceDriven2DBranchless.cu:73,423,513); with prefix sums evaluated
// exactly they equal the box-integral projector with that switch.
static void hard_switch(std::vector<DDView>& v) {
    for (auto& o : v) {
        const bool ymain = (o.flags & 4) != 0;
        o.wx = ymain ? 1.f : 0.f;
        o.flags = (o.flags & 4) | (ymain ? 1 : 2);
    }
}

// ---------------------------------------------------------------------------
// layout transforms
// ---------------------------------------------------------------------------
// planes P[c][y][x] (c < C)  ->  O[major * rs + minor * cs + c] with (major,
// minor) = (y, x) (xmajor=false) or (x, y) (xmajor=true); rs >= minor extent *
// cs is the row stride; channels >= C are written as zero.
// Block (bx, by, bz) = (32 pixels of a row, 32 channels, row y) of the copy.
__device__ inline void to_channel_last_block(float* __restrict__ out, const float* __restrict__ in, int C, int cs,
                                             int nx, int ny, bool xmajor, size_t rs, int bx, int by, int bz) {
    __shared__ float tile[32][33];
    const int x0 = bx * 32, c0 = by * 32, y = bz;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, x = x0 + tx;
        tile[k][tx] = (c < C && x < nx) ? in[((size_t)c * ny + y) * nx + x] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int x = x0 + k, c = c0 + tx;
        if (x < nx && c < cs) {
            const size_t at = xmajor ? (size_t)x * rs + (size_t)y * cs : (size_t)y * rs + (size_t)x * cs;
            out[at + c] = tile[tx][k];
        }
    }
}

__global__ __launch_bounds__(256) void to_channel_last(float* __restrict__ out, const float* __restrict__ in, int C,
                                                       int cs, int nx, int ny, bool xmajor, size_t rs) {
    to_channel_last_block(out, in, C, cs, nx, ny, xmajor, rs, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// forward projection
// ---------------------------------------------------------------------------
struct FPRow {        // one output row (b, iv) of a channel chunk
    int b, iv;
    int j0, j1;       // accumulator indices of slices iz and iz1 relative to the chunk
    float w1;         // weight of iz1 (iz gets 1 - w1)
    float pad[3];
};

// Footprint of one line [x1, x2] (index units, unclamped) on a pixel row of
// length n: the reference's IntegralBoxX (cuda/cudaMath.cu:533-565) divided by
// (x2 - x1), applied to a channel vector at 'row' (stride cs between pixels).
// XC: the channel after the chunk (the next chunk's first slice, for the z
// interpolation across chunks) is accumulated too.
template <int CH, bool XC>
__device__ inline void dd_line(float* acc, const float* __restrict__ row, int cs, float x1, float x2, int n, float scale) {
    const float lo = min_f(x1, x2), hi = max_f(x1, x2);
    const float a = med_clamp(lo, 0.f, (float)n), b = med_clamp(hi, 0.f, (float)n);
    if (!(b > a)) return;
    const float inv = fast_div(scale, hi - lo);
    const int i1 = min((int)a, n - 1), i2 = min((int)b, n - 1);
    for (int i = i1; i <= i2; i++) {
        const float w = (fminf(b, (float)(i + 1)) - fmaxf(a, (float)i)) * inv;
        const float* p = row + (size_t)i * cs;
        if constexpr (CH % 4 == 0) {
#pragma unroll
            for (int j = 0; j < CH; j += 4) {
                const float4 q = *reinterpret_cast<const float4*>(p + j);
                acc[j] = fmaf(w, q.x, acc[j]);
                acc[j + 1] = fmaf(w, q.y, acc[j + 1]);
                acc[j + 2] = fmaf(w, q.z, acc[j + 2]);
                acc[j + 3] = fmaf(w, q.w, acc[j + 3]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) acc[j] = fmaf(w, p[j], acc[j]);
        }
        if constexpr (XC) acc[CH] = fmaf(w, p[CH], acc[CH]);
    }
}

// The two edge rays of detector column iu in image index coordinates: from
// s1 to d1 and from s2 to d2 (cuda/distanceDriven2DBoxInt.cu:141-158 parallel
// with its virtual source distance, :211-226 fan).  Returns the column's fan
// angle (0 for parallel beams).  The rotations and sums are written as
// explicit fused multiply-adds, so that every kernel inlining this function
// rounds them alike (the compiler's own contraction of a*b + c*d may fuse
// either product, depending on the surrounding code).
template <bool FAN>
__device__ inline float fp_cell_edges(int iu, const DDView& V, const Grid& g, const Det& d, float dsd, float dso,
                                      float2& s1, float2& s2, float2& d1, float2& d2) {
    const float c = V.c, s = V.s;
    float a = 0.f;
    auto rot = [&](float px, float py) { return make_float2(fmaf(px, c, -(py * s)), fmaf(px, s, py * c)); };
    if (FAN) {
        a = (-((float)iu - (float)(d.nu - 1) / 2.0f) - d.off_u) * d.du;
        const float2 src = make_float2(dso * s, -dso * c);
        const float e1x = dsd * sinf(a - d.du / 2), e1y = fmaf(dsd, cosf(a - d.du / 2), -dso);
        const float e2x = dsd * sinf(a + d.du / 2), e2y = fmaf(dsd, cosf(a + d.du / 2), -dso);
        s1 = s2 = src;
        d1 = rot(e1x, e1y);
        d2 = rot(e2x, e2y);
    } else {
        const float u = (-((float)iu - (float)(d.nu - 1) / 2.0f) - d.off_u) * d.du;
        const float vd = fmaf((float)g.nx, g.dx, (float)g.ny * g.dy);
        s1 = rot(u - d.du / 2, -vd);
        s2 = rot(u + d.du / 2, -vd);
        d1 = rot(u - d.du / 2, vd);
        d2 = rot(u + d.du / 2, vd);
    }
    auto toimg = [&](float2 p) {
        return make_float2((p.x - g.cx) / g.dx + g.nx / 2.0f, (p.y - g.cy) / g.dy + g.ny / 2.0f);
    };
    s1 = toimg(s1); s2 = toimg(s2); d1 = toimg(d1); d2 = toimg(d2);
    return a;
}

// A block is 128 lanes: 128 >> lseg detector columns times 2^lseg segments of
// the image rows (lseg 0 or 1: one segment per wavefront, so the row loop
// stays scalar).  Each lane integrates its column's footprint over its
// segment; the segments' accumulators are summed in LDS in a fixed order.
// Splitting the rows doubles the waves of launches with few columns, views
// and channel chunks (2D slices) so that they can hide the row gathers'
// latency.
// VPB > 1: a block is VPB wavefronts of VPB consecutive views over the same
// 64 columns (one segment): the views' footprints on a row nearly coincide, so
// the wavefronts of a block read the same image lines at about the same time
// and three of four find them in the CU's L1 (measured at 512x512x32, 984
// views: 6.7 ms with one view per block, 6.7 ms with two, 5.8 ms with four).
// Block (bx, by, bz) = (column block, view or view group, channel chunk).
template <int CH, bool FAN, bool XC, int VPB>
__device__ inline void dd_fp_block(
        float* __restrict__ prj, const float* __restrict__ imgY, const float* __restrict__ imgX,
        const DDView* __restrict__ views, const FPRow* __restrict__ rows, const int* __restrict__ row_start, int cs,
        const Grid& g, const Det& d, int nview, float dsd, float dso, bool force, int lseg, int bx, int by, int bz) {
    __shared__ float lds[VPB > 1 ? 64 * VPB : 128][CH + 2];
    const int lcol = VPB > 1 ? 6 : 7 - lseg;
    const int col = threadIdx.x & ((1 << lcol) - 1);
    const int seg = VPB > 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> lcol);   // uniform per wavefront
    const int nseg = VPB > 1 ? 1 : 1 << lseg;
    const int iu = (bx << lcol) + col;
    const int view = VPB > 1 ? by * VPB + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : by;
    const int chunk = bz;
    if (VPB > 1 && view >= nview) return;   // no barriers in this form
    const int c0 = chunk * CH;
    const DDView V = views[view];
    float acc[CH + 1];
#pragma unroll
    for (int j = 0; j <= CH; j++) acc[j] = 0.f;
    const bool active = iu < d.nu;
    if (active) {
        float2 s1, s2, d1, d2;
        const float a = fp_cell_edges<FAN>(iu, V, g, d, dsd, dso, s1, s2, d1, d2);
        const float* __restrict__ bY = imgY + c0;
        const float* __restrict__ bX = imgX + c0;
        if (V.flags & 1) {
            // trace along y, cuda/distanceDriven2DBoxInt.cu:48-71
            const float r1 = (d1.x - s1.x) / (d1.y - s1.y);
            const float r2 = (d2.x - s2.x) / (d2.y - s2.y);
            const float norm = V.wx * g.dy / fabsf(cosf(V.deg - a));
            // rows addressed from the int extents (a 64-bit row-stride argument
            // instead measured 8% slower at 512x512x32)
            const int ye = (int)(((long)g.ny * (seg + 1)) >> lseg);
            for (int iy = (int)(((long)g.ny * seg) >> lseg); iy < ye; iy++) {
                const float yy = (float)iy + 0.5f;
                const float x1 = fmaf(r1, yy - s1.y, s1.x);
                const float x2 = fmaf(r2, yy - s2.y, s2.x);
                dd_line<CH, XC>(acc, bY + (size_t)iy * g.nx * cs, cs, x1, x2, g.nx, norm);
            }
        }
        if (V.flags & 2) {
            // trace along x, cuda/distanceDriven2DBoxInt.cu:74-98
            const float r1 = (d1.y - s1.y) / (d1.x - s1.x);
            const float r2 = (d2.y - s2.y) / (d2.x - s2.x);
            const float norm = (1.f - V.wx) * g.dx / fabsf(sinf(V.deg - a));
            const int xe = (int)(((long)g.nx * (seg + 1)) >> lseg);
            for (int ix = (int)(((long)g.nx * seg) >> lseg); ix < xe; ix++) {
                const float xx = (float)ix + 0.5f;
                const float y1 = fmaf(r1, xx - s1.x, s1.y);
                const float y2 = fmaf(r2, xx - s2.x, s2.y);
                dd_line<CH, XC>(acc, bX + (size_t)ix * g.ny * cs, cs, y1, y2, g.ny, norm);
            }
        }
    }
#pragma unroll
    for (int j = 0; j <= CH; j++) lds[threadIdx.x][j] = acc[j];
    if (nseg > 1) {
        __syncthreads();
        if (seg != 0) return;
        // segment 0 sums the other segments' accumulators into its own row
#pragma unroll
        for (int j = 0; j <= CH; j++) {
            float v = acc[j];
            for (int s = 1; s < nseg; s++) v += lds[(s << lcol) + col][j];
            lds[col][j] = v;
        }
    }
    // rows are read back by the owning lane only: no further barrier
    if (!active) return;
    const float rev = (!FAN && force) ? V.rev : 1.f;
    const int r0 = row_start[chunk], r1 = row_start[chunk + 1];
    const int mine = VPB > 1 ? threadIdx.x : col;
    for (int r = r0; r < r1; r++) {
        const FPRow R = rows[r];
        const float val = (1.f - R.w1) * lds[mine][R.j0] + R.w1 * lds[mine][R.j1];
        prj[(((size_t)R.b * nview + view) * d.nv + R.iv) * d.nu + iu] = val * rev;
    }
}

template <int CH, bool FAN, bool XC, int VPB>
__global__ __launch_bounds__(VPB > 1 ? 64 * VPB : 128) void dd_fp_kernel(
        float* __restrict__ prj, const float* __restrict__ imgY, const float* __restrict__ imgX,
        const DDView* __restrict__ views, const FPRow* __restrict__ rows, const int* __restrict__ row_start, int cs,
        Grid g, Det d, int nview, float dsd, float dso, bool force, int lseg) {
    dd_fp_block<CH, FAN, XC, VPB>(prj, imgY, imgX, views, rows, row_start, cs, g, d, nview, dsd, dso, force, lseg,
                                  blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// Forward projection, several lanes per detector column
// ---------------------------------------------------------------------------
// The kernel above gives each lane a whole channel vector, so one float4 load
// of a wavefront touches a different cache line for nearly every lane (64
// columns cover ~40 pixels of a row) and uses 16 bytes of each: the L1's line
// lookups, not HBM, set its speed.  Here LPC lanes share a column and split
// its CH channels (lane s of the column holds the float4s at channel offsets
// 4 (s + LPC k)), so load k of a wavefront reads one contiguous 16 LPC-byte
// piece of each pixel it touches, and 64 / LPC columns touch far fewer
// pixels.  The footprints of LPC consecutive image rows are evaluated one per
// lane and broadcast within the column's lanes.  Pixels are CH floats (no padding), one chunk, no channel after
// it (channel CH reads as 0).  A block is 4 wavefronts = 4 consecutive views
// over the same 64 / LPC columns, for the L1 sharing described above.
// Pixels i1..i2 of a row weighted by their overlap with [a, b] times inv,
// applied to NQ float4s 4 LPC floats apart: the lane's floats of pixel i start
// at row[i cs + lane_off].  'row' is the same for the whole wavefront, so
// the loads take it as a scalar base plus a 32-bit per-lane byte offset.  The
// pixel edges i and i + 1 are carried as floats (exact below 2^24);
// min(x, y) = med3(x, y, -3e38) and max(x, y) = med3(x, y, 3e38) for these
// finite operands, without re-canonicalising the loop-invariant bounds every
// pixel as fminf / fmaxf do.
template <int NQ, int LPC>
__device__ inline void split_pixels(float* acc, const float* __restrict__ row, int lane_off, int cs, float a, float b,
                                    float inv, int i1, int i2) {
    if (i1 > i2) return;
    const char* __restrict__ rb = reinterpret_cast<const char*>(row);
    unsigned off = (unsigned)(i1 * cs + lane_off) * 4u;
    const unsigned step = (unsigned)cs * 4u;
    float fi = (float)i1;
    const float fe = (float)i2;
    for (;;) {
        const float fn = fi + 1.f;
        const float w = (__builtin_amdgcn_fmed3f(b, fn, -3.0e38f) - __builtin_amdgcn_fmed3f(a, fi, 3.0e38f)) * inv;
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            const float4 q = *reinterpret_cast<const float4*>(rb + off + 16u * LPC * k);
            acc[4 * k] = fmaf(w, q.x, acc[4 * k]);
            acc[4 * k + 1] = fmaf(w, q.y, acc[4 * k + 1]);
            acc[4 * k + 2] = fmaf(w, q.z, acc[4 * k + 2]);
            acc[4 * k + 3] = fmaf(w, q.w, acc[4 * k + 3]);
        }
        if (!(fn <= fe)) break;
        fi = fn;
        off += step;
    }
}

template <int NQ, int LPC>
__device__ inline void dd_line_split(float* acc, const float* __restrict__ row, int lane_off, int cs, float x1,
                                     float x2, int n, float scale) {
    const float lo = min_f(x1, x2), hi = max_f(x1, x2);
    const float a = med_clamp(lo, 0.f, (float)n), b = med_clamp(hi, 0.f, (float)n);
    if (!(b > a)) return;
    const float inv = fast_div(scale, hi - lo);
    split_pixels<NQ, LPC>(acc, row, lane_off, cs, a, b, inv, min((int)a, n - 1), min((int)b, n - 1));
}

// One trace of a column for dd_fp_split_kernel: rows r = 0 .. nrows-1 of the
// channel-last copy at 'img' (row stride rstride floats, pixel stride CH),
// the column's lines at row coordinate r + 0.5 being fmaf(r1, t - o1, q1) and
// fmaf(r2, t - o2, q2) on rows of n pixels.  The footprint of row t0 + s
// (clamped ends, 1 / width, first and last pixel: dd_line_split's arithmetic)
// is evaluated by lane s of the column's LPC lanes and broadcast to the others.
template <int NQ, int LPC>
__device__ inline void fp_trace_split(float* acc, const float* __restrict__ img, size_t rstride, int nrows, int n,
                                      float r1, float o1, float q1, float r2, float o2, float q2, float norm, int sub,
                                      int lead) {
    constexpr int CH = 4 * LPC * NQ;
    for (int t0 = 0; t0 < nrows; t0 += LPC) {
        float ma = 0.f, mb = 0.f, minv = 0.f;
        int mi1 = 0, mi2 = -1;   // empty
        if (t0 + sub < nrows) {
            const float tt = (float)(t0 + sub) + 0.5f;
            const float e1 = fmaf(r1, tt - o1, q1), e2 = fmaf(r2, tt - o2, q2);
            const float x1 = min_f(e1, e2), x2 = max_f(e1, e2);
            const float a = med_clamp(x1, 0.f, (float)n), b = med_clamp(x2, 0.f, (float)n);
            if (b > a) {
                ma = a; mb = b;
                minv = fast_div(norm, x2 - x1);
                mi1 = min((int)a, n - 1);
                mi2 = min((int)b, n - 1);
            }
        }
#pragma unroll
        for (int t = 0; t < LPC; t++) {
            if (t0 + t >= nrows) break;   // uniform
            const float a = group_bcast<LPC>(ma, t, lead), b = group_bcast<LPC>(mb, t, lead);
            const float inv = group_bcast<LPC>(minv, t, lead);
            const int i1 = group_bcast<LPC>(mi1, t, lead), i2 = group_bcast<LPC>(mi2, t, lead);
            split_pixels<NQ, LPC>(acc, img + (size_t)(t0 + t) * rstride, 4 * sub, CH, a, b, inv, i1, i2);
        }
    }
}

// Block (bx, by) = (64 / LPC detector columns, 4 consecutive views).
template <int CH, int LPC, bool FAN>
__device__ inline void dd_fp_split_block(
        float* __restrict__ prj, const float* __restrict__ imgY, const float* __restrict__ imgX,
        const DDView* __restrict__ views, const FPRow* __restrict__ rows, int nrow, const Grid& g, const Det& d,
        int nview, float dsd, float dso, bool force, int bx, int by) {
    static_assert(CH % (4 * LPC) == 0, "a lane holds whole float4s");
    constexpr int NQ = CH / (4 * LPC);   // float4s per lane and pixel
    constexpr int NC = 64 / LPC;         // columns per wavefront
    __shared__ float lds[4][NC][CH + 1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int view = by * 4 + wave;
    if (view >= nview) return;   // every exchange below stays within the wavefront
    const int c = lane / LPC, sub = lane % LPC;
    const int iu = bx * NC + c;
    const DDView V = views[view];
    float acc[4 * NQ];
#pragma unroll
    for (int j = 0; j < 4 * NQ; j++) acc[j] = 0.f;
    if (iu < d.nu) {
        float2 s1, s2, d1, d2;
        const float a = fp_cell_edges<FAN>(iu, V, g, d, dsd, dso, s1, s2, d1, d2);
        if (V.flags & 1) {
            // trace along y, cuda/distanceDriven2DBoxInt.cu:48-71
            const float r1 = (d1.x - s1.x) / (d1.y - s1.y);
            const float r2 = (d2.x - s2.x) / (d2.y - s2.y);
            const float norm = V.wx * g.dy / fabsf(cosf(V.deg - a));
            fp_trace_split<NQ, LPC>(acc, imgY, (size_t)g.nx * CH, g.ny, g.nx, r1, s1.y, s1.x, r2, s2.y,
                                    s2.x, norm, sub, lane - sub);
        }
        if (V.flags & 2) {
            // trace along x, cuda/distanceDriven2DBoxInt.cu:74-98
            const float r1 = (d1.y - s1.y) / (d1.x - s1.x);
            const float r2 = (d2.y - s2.y) / (d2.x - s2.x);
            const float norm = (1.f - V.wx) * g.dx / fabsf(sinf(V.deg - a));
            fp_trace_split<NQ, LPC>(acc, imgX, (size_t)g.ny * CH, g.nx, g.ny, r1, s1.x, s1.y, r2, s2.x,
                                    s2.y, norm, sub, lane - sub);
        }
    }
    // gather the column's channel vector in LDS (channel CH is 0) and read
    // back the detector rows: lane = (column lane % NC, row group lane / NC)
    float* col = &lds[wave][c][0];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
#pragma unroll
        for (int e = 0; e < 4; e++) col[4 * (sub + LPC * k) + e] = acc[4 * k + e];
    }
    if (sub == 0) col[CH] = 0.f;
    // the LDS writes above are complete before any lane of the wavefront reads
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    const int oc = lane % NC, rg = lane / NC;
    const int ou = bx * NC + oc;
    if (ou >= d.nu) return;
    const float rev = (!FAN && force) ? V.rev : 1.f;
    const float* mine = &lds[wave][oc][0];
    for (int r = rg; r < nrow; r += LPC) {
        const FPRow R = rows[r];
        const float val = (1.f - R.w1) * mine[R.j0] + R.w1 * mine[R.j1];
        prj[(((size_t)R.b * nview + view) * d.nv + R.iv) * d.nu + ou] = val * rev;
    }
}

template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void dd_fp_split_kernel(
        float* __restrict__ prj, const float* __restrict__ imgY, const float* __restrict__ imgX,
        const DDView* __restrict__ views, const FPRow* __restrict__ rows, int nrow, Grid g, Det d, int nview,
        float dsd, float dso, bool force) {
    dd_fp_split_block<CH, LPC, FAN>(prj, imgY, imgX, views, rows, nrow, g, d, nview, dsd, dso, force, blockIdx.x,
                                    blockIdx.y);
}

// ---------------------------------------------------------------------------
// Forward projection with LDS-staged image rows
// ---------------------------------------------------------------------------
// A block is VB views x 64 detector columns (one wavefront per view; the views
// of a block share their trace flags).  For each image row of a trace, the
// union of the block's footprints on that row lies between the outermost cell
// edges of the block (rays from one source do not cross between source and
// detector), so the block copies that pixel range of the row (channel-last,
// CHP floats per pixel) into LDS once and every lane integrates its box from
// LDS: the row is fetched once per VB views instead of once per view, and the
// per-lane reads move from the vector-memory path to LDS.  Rows are double
// buffered with one barrier per row.
constexpr int kFPViews = 4;   // views per block (VB)

struct FPGroup {               // consecutive views with equal trace flags
    int first, count;
};

// Integral of one lane's line pair over staged pixels [xlo, xlo + width) held
// in LDS at 'buf' (CHP floats per pixel); the same arithmetic as dd_line.
template <int CH, int CHP>
__device__ inline void dd_line_lds(float* acc, const float* buf, int xlo, float x1, float x2, int n, float scale) {
    const float lo = min_f(x1, x2), hi = max_f(x1, x2);
    const float a = med_clamp(lo, 0.f, (float)n), b = med_clamp(hi, 0.f, (float)n);
    if (!(b > a)) return;
    const float inv = fast_div(scale, hi - lo);
    const int i1 = min((int)a, n - 1), i2 = min((int)b, n - 1);
    for (int i = i1; i <= i2; i++) {
        const float w = (fminf(b, (float)(i + 1)) - fmaxf(a, (float)i)) * inv;
        const float* p = buf + (i - xlo) * CHP;
#pragma unroll
        for (int j = 0; j < CHP; j += 4) {
            const float4 q = *reinterpret_cast<const float4*>(p + j);
            if (j < CH + 1) acc[j] = fmaf(w, q.x, acc[j]);
            if (j + 1 < CH + 1) acc[j + 1] = fmaf(w, q.y, acc[j + 1]);
            if (j + 2 < CH + 1) acc[j + 2] = fmaf(w, q.z, acc[j + 2]);
            if (j + 3 < CH + 1) acc[j + 3] = fmaf(w, q.w, acc[j + 3]);
        }
    }
}

// One trace (rows along y or along x) of a block.  The staged pixel range of
// each row is copied from global memory straight into LDS by the texture path's
// LDS DMA (global_load_lds_dwordx4: each wavefront instruction fills 1 KiB of
// LDS, no registers in between), double buffered: the next row's copy is in
// flight while the block integrates the current row from LDS.
template <int CH, int CHP, int VB>
__device__ inline void fp_trace_lds(float* acc, float* stage, const float (*edges)[4][3], int nviews_blk,
                                    const float* __restrict__ src, int cs, long rs, int c0, int nrow, int nlen, int wmax,
                                    bool active, float r1, float o1, float b1, float r2, float o2, float b2,
                                    float norm) {
    constexpr int Q = CHP / 4;             // float4 per pixel
    constexpr int nthr = 64 * VB;
    const int wave = threadIdx.x >> 6;
    // pixel range of row 'row' staged by the block (the same for every lane)
    auto range = [&](int row, int& lo, int& width) {
        const float t = (float)row + 0.5f;
        float mn = 3.0e38f, mx = -3.0e38f;
        for (int w = 0; w < nviews_blk; w++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float x = fmaf(edges[w][k][0], t - edges[w][k][1], edges[w][k][2]);
                mn = fminf(mn, x);
                mx = fmaxf(mx, x);
            }
        // clamp both ends into [-1, nlen + 1] (fminf/fmaxf also map NaN there)
        mn = fminf(fmaxf(mn, -1.f), (float)nlen + 1.f);
        mx = fmaxf(fminf(mx, (float)nlen + 1.f), -1.f);
        lo = __builtin_amdgcn_readfirstlane(max(0, (int)floorf(mn) - 1));
        const int hi = min(nlen - 1, (int)floorf(mx) + 1);
        width = __builtin_amdgcn_readfirstlane(min(max(hi - lo + 1, 0), wmax));
    };
    // copy row 'row' (pixels [lo, lo + width), CHP floats each) into buf:
    // item t (pixel t / Q, float4 t % Q) lands at buf[4 t]
    auto fetch = [&](int row, float* buf, int& lo) {
        int width;
        range(row, lo, width);
        const float* base = src + (size_t)row * rs + (size_t)lo * cs + c0;
        const int total = width * Q;
        for (int k0 = 0; k0 < total; k0 += nthr) {
            const int t = k0 + threadIdx.x;
            if (t < total) {
                const int px = t / Q, q = t - px * Q;
                __builtin_amdgcn_global_load_lds(
                    (const void*)(base + (size_t)px * cs + 4 * q),
                    (__attribute__((address_space(3))) void*)(buf + 4 * (k0 + 64 * wave)), 16, 0, 0);
            }
        }
    };
    int lo_cur = 0, lo_next = 0;
    fetch(0, stage, lo_cur);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int row = 0; row < nrow; row++) {
        float* cur = stage + (row & 1) * wmax * CHP;
        float* nxt = stage + ((row + 1) & 1) * wmax * CHP;
        if (row + 1 < nrow) fetch(row + 1, nxt, lo_next);
        if (active) {
            const float t = (float)row + 0.5f;
            const float x1 = fmaf(r1, t - o1, b1);
            const float x2 = fmaf(r2, t - o2, b2);
            dd_line_lds<CH, CHP>(acc, cur, lo_cur, x1, x2, nlen, norm);
        }
        lo_cur = lo_next;
        // the next row has landed in LDS and every wave is done with this one
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
}

template <int CH, bool FAN>
__global__ __launch_bounds__(64 * kFPViews) void dd_fp_lds_kernel(
        float* __restrict__ prj, const float* __restrict__ imgY, const float* __restrict__ imgX,
        const DDView* __restrict__ views, const FPGroup* __restrict__ groups, const FPRow* __restrict__ rows,
        const int* __restrict__ row_start, int cs, long rsY, long rsX, int wmax, Grid g, Det d, int nview, float dsd,
        float dso, bool force) {
    constexpr int VB = kFPViews;
    constexpr int CHP = ((CH + 1) + 3) / 4 * 4;
    extern __shared__ float4 smem4[];
    float* stage = reinterpret_cast<float*>(smem4);
    __shared__ float edges[VB][4][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int col0 = blockIdx.x * 64;
    const FPGroup G = groups[blockIdx.y];
    const int chunk = blockIdx.z, c0 = chunk * CH;
    const bool vok = w < G.count;
    const int view = G.first + (vok ? w : 0);
    const DDView V = views[view];
    const int flags = views[G.first].flags;
    const int iu = col0 + lane;
    const bool active = vok && iu < d.nu;
    const int last = min(63, d.nu - 1 - col0);
    float acc[CH + 1];
#pragma unroll
    for (int j = 0; j <= CH; j++) acc[j] = 0.f;
    float2 s1, s2, d1, d2;
    const float a = fp_cell_edges<FAN>(min(iu, d.nu - 1), V, g, d, dsd, dso, s1, s2, d1, d2);
    if (flags & 1) {
        // trace along y, cuda/distanceDriven2DBoxInt.cu:48-71
        const float r1 = (d1.x - s1.x) / (d1.y - s1.y);
        const float r2 = (d2.x - s2.x) / (d2.y - s2.y);
        if (vok && (lane == 0 || lane == last)) {
            const int k = lane == 0 ? 0 : 2;
            edges[w][k][0] = r1; edges[w][k][1] = s1.y; edges[w][k][2] = s1.x;
            edges[w][k + 1][0] = r2; edges[w][k + 1][1] = s2.y; edges[w][k + 1][2] = s2.x;
        }
        __syncthreads();
        const float norm = V.wx * g.dy / fabsf(cosf(V.deg - a));
        fp_trace_lds<CH, CHP, VB>(acc, stage, edges, G.count, imgY, cs, rsY, c0, g.ny, g.nx, wmax, active,
                                         r1, s1.y, s1.x, r2, s2.y, s2.x, norm);
    }
    if (flags & 2) {
        // trace along x, cuda/distanceDriven2DBoxInt.cu:74-98
        const float q1 = (d1.y - s1.y) / (d1.x - s1.x);
        const float q2 = (d2.y - s2.y) / (d2.x - s2.x);
        __syncthreads();                   // edges of a previous trace are no longer read
        if (vok && (lane == 0 || lane == last)) {
            const int k = lane == 0 ? 0 : 2;
            edges[w][k][0] = q1; edges[w][k][1] = s1.x; edges[w][k][2] = s1.y;
            edges[w][k + 1][0] = q2; edges[w][k + 1][1] = s2.x; edges[w][k + 1][2] = s2.y;
        }
        __syncthreads();
        const float norm = (1.f - V.wx) * g.dx / fabsf(sinf(V.deg - a));
        fp_trace_lds<CH, CHP, VB>(acc, stage, edges, G.count, imgX, cs, rsX, c0, g.nx, g.ny, wmax, active,
                                        q1, s1.x, s1.y, q2, s2.x, s2.y, norm);
    }
    // z interpolation of the detector rows from the accumulators (each lane
    // reads back only its own accumulators)
    float* mine = stage + threadIdx.x * (CH + 2);
#pragma unroll
    for (int j = 0; j <= CH; j++) mine[j] = acc[j];
    if (!active) return;
    const float rev = (!FAN && force) ? V.rev : 1.f;
    const int r0 = row_start[chunk], rend = row_start[chunk + 1];
    for (int r = r0; r < rend; r++) {
        const FPRow R = rows[r];
        const float val = (1.f - R.w1) * mine[R.j0] + R.w1 * mine[R.j1];
        prj[(((size_t)R.b * nview + view) * d.nv + R.iv) * d.nu + iu] = val * rev;
    }
}

// Host copy of fp_cell_edges in FP64: the widest staged range (pixels) of any
// block, or a huge value when a ray runs (nearly) along a traced row.
static double fp_stage_width(const std::vector<DDView>& views, const std::vector<FPGroup>& groups, const Grid& g,
                             const Det& d, float dsd, float dso, bool fan) {
    double wmax = 0;
    auto edge_lines = [&](int iu, const DDView& V, double ln[2][2][3]) {
        // ln[trace][edge] = (slope, origin, offset): x = slope * (t - origin) + offset
        const double c = V.c, s = V.s;
        double s1[2], s2[2], d1[2], d2[2];
        if (fan) {
            const double a = (-((double)iu - (d.nu - 1) / 2.0) - d.off_u) * d.du;
            const double src[2] = {dso * s, -dso * c};
            const double e1x = dsd * std::sin(a - d.du / 2), e1y = -dso + dsd * std::cos(a - d.du / 2);
            const double e2x = dsd * std::sin(a + d.du / 2), e2y = -dso + dsd * std::cos(a + d.du / 2);
            s1[0] = s2[0] = src[0]; s1[1] = s2[1] = src[1];
            d1[0] = e1x * c - e1y * s; d1[1] = e1x * s + e1y * c;
            d2[0] = e2x * c - e2y * s; d2[1] = e2x * s + e2y * c;
        } else {
            const double u = (-((double)iu - (d.nu - 1) / 2.0) - d.off_u) * d.du;
            const double vd = (double)g.nx * g.dx + (double)g.ny * g.dy;
            auto rot = [&](double px, double py, double* o) { o[0] = px * c - py * s; o[1] = px * s + py * c; };
            rot(u - d.du / 2, -vd, s1); rot(u + d.du / 2, -vd, s2);
            rot(u - d.du / 2, vd, d1); rot(u + d.du / 2, vd, d2);
        }
        auto toimg = [&](double* p) { p[0] = (p[0] - g.cx) / g.dx + g.nx / 2.0; p[1] = (p[1] - g.cy) / g.dy + g.ny / 2.0; };
        toimg(s1); toimg(s2); toimg(d1); toimg(d2);
        ln[0][0][0] = (d1[0] - s1[0]) / (d1[1] - s1[1]); ln[0][0][1] = s1[1]; ln[0][0][2] = s1[0];
        ln[0][1][0] = (d2[0] - s2[0]) / (d2[1] - s2[1]); ln[0][1][1] = s2[1]; ln[0][1][2] = s2[0];
        ln[1][0][0] = (d1[1] - s1[1]) / (d1[0] - s1[0]); ln[1][0][1] = s1[0]; ln[1][0][2] = s1[1];
        ln[1][1][0] = (d2[1] - s2[1]) / (d2[0] - s2[0]); ln[1][1][1] = s2[0]; ln[1][1][2] = s2[1];
    };
    const int ncb = ceil_div(d.nu, 64);
    for (const FPGroup& G : groups)
        for (int cb = 0; cb < ncb; cb++) {
            const int cols[2] = {cb * 64, std::min(d.nu - 1, cb * 64 + 63)};
            for (int tr = 0; tr < 2; tr++) {
                if (!(views[G.first].flags & (1 << tr))) continue;
                const int nrow = tr == 0 ? g.ny : g.nx;
                const double nlen = tr == 0 ? g.nx : g.ny;
                // the outermost edge positions are a convex (max) and a concave
                // (min) function of the row, so their extremes over the rows
                // are at the two end rows; the kernel clamps both ends into
                // [-1, nlen + 1], so the staged width is at most
                // max over the ends of the clamped max - min of the clamped min
                double hi = -1e300, lo = 1e300;
                for (int end = 0; end < 2; end++) {
                    const double t = end == 0 ? 0.5 : nrow - 0.5;
                    double mn = 1e300, mx = -1e300;
                    for (int k = 0; k < G.count; k++)
                        for (int ci = 0; ci < 2; ci++) {
                            double ln[2][2][3];
                            edge_lines(cols[ci], views[G.first + k], ln);
                            for (int e = 0; e < 2; e++) {
                                const double x = ln[tr][e][0] * (t - ln[tr][e][1]) + ln[tr][e][2];
                                if (!std::isfinite(x)) return 1e300;
                                mn = std::min(mn, x);
                                mx = std::max(mx, x);
                            }
                        }
                    hi = std::max(hi, std::min(std::max(mx, -1.0), nlen + 1.0));
                    lo = std::min(lo, std::min(std::max(mn, -1.0), nlen + 1.0));
                }
                wmax = std::max(wmax, hi - lo);
            }
        }
    return wmax;
}

static int pick_ch(int C) {
    int ch = 1;
    while (ch < C && ch < 32) ch *= 2;
    return ch;
}

// Geometry-only launch data of the forward projector for one batch count (host
// side: uploaded per call by dd2d_fp, once by a DDPlan).
struct DDFPSetup {
    int C = 0, CH = 1, nchunk = 1, cs = 0;
    std::vector<FPRow> rows;
    std::vector<int> start;
    std::vector<FPGroup> groups;
    bool use_lds = false, needY = false, needX = false, force = false;
    int wmax = 0;
    int lseg = 0;   // direct kernel: 2^lseg row segments per column
    int lpc = 1;    // lanes per detector column (> 1: dd_fp_split_kernel)
    bool xc = true; // pixels carry the channel after the chunk
    long rsY = 0, rsX = 0;   // row strides (floats) of the y-major and x-major copies (n * cs)
};

// allow_lds: the LDS-staged form may be chosen (by the dd_fp_mode tuning key);
// lpc_req > 0: lanes per column requested instead of the tuning key, reduced
// to what the chunk allows (CH / 4; 1 = the one-lane kernel).
static DDFPSetup dd_fp_setup(const std::vector<DDView>& views, int B, const Grid& g, const Det& d, int nview,
                             float dsd, float dso, int type, bool fan, bool allow_lds = true, int lpc_req = 0) {
    const bool branchless = (type & DD_BRANCHLESS) != 0;
    DDFPSetup s;
    s.C = B * g.nz;
    s.CH = pick_ch(s.C);
    s.nchunk = ceil_div(s.C, s.CH);
    const int CH = s.CH, nchunk = s.nchunk;
    // rows of each chunk: detector row iv of batch b belongs to the chunk holding
    // channel b*nz + iz(iv); the z mapping follows cuda/distanceDriven2DBoxInt.cu:35-42,214,227
    std::vector<std::vector<FPRow>> per(nchunk);
    for (int b = 0; b < B; b++)
        for (int iv = 0; iv < d.nv; iv++) {
            const double zr = ((double)iv - (d.nv - 1) / 2.0 + d.off_v) * d.dv;
            double z = (zr - g.cz) / g.dz + g.nz / 2.0 - 0.5;
            z = std::min((double)g.nz, std::max(0.0, z));
            int iz = (int)z, iz1 = iz + 1;
            iz = clampi(iz, 0, g.nz);
            iz1 = clampi(iz1, 0, g.nz);
            const double wz = z - iz;
            const int cl = b * g.nz + iz;
            const int k = cl / CH;
            FPRow r{};
            r.b = b; r.iv = iv;
            r.j0 = cl - k * CH;
            r.j1 = b * g.nz + iz1 - k * CH;
            r.w1 = (float)wz;
            per[k].push_back(r);
        }
    s.start.assign(nchunk + 1, 0);
    for (int k = 0; k < nchunk; k++) {
        s.start[k] = (int)s.rows.size();
        s.rows.insert(s.rows.end(), per[k].begin(), per[k].end());
    }
    s.start[nchunk] = (int)s.rows.size();

    // LDS-staged kernel: groups of up to kFPViews consecutive views with equal
    // trace flags; staged width bounded in FP64 (+ margins for the +-1 pixel
    // guard and floor rounding)
    for (int i = 0; i < nview;) {
        int n = 1;
        while (n < kFPViews && i + n < nview && (views[i + n].flags & 3) == (views[i].flags & 3)) n++;
        s.groups.push_back(FPGroup{i, n});
        i += n;
    }
    const int CHP = ((CH + 1) + 3) / 4 * 4;
    const int wcap = 65536 / (2 * CHP * 4);   // two row buffers in 64 KiB of LDS
    // the direct kernel is the default (measured faster at 512x512x32: the
    // staged kernel's prefetch registers spill)
    const bool want_lds = allow_lds && tuning_get("dd_fp_mode", 1) == 2 && s.groups.size() <= 65535;
    const double wneed = want_lds ? fp_stage_width(views, s.groups, g, d, dsd, dso, fan) + 6.0 : 1e300;
    s.use_lds = want_lds && wneed <= (double)wcap;
    s.wmax = s.use_lds ? (int)std::ceil(wneed) : 0;
    // channel-last pixel stride: with one chunk the direct kernel needs no
    // channel after the chunk (rows interpolate between slices of the chunk),
    // so a pixel is CH floats (one 128-byte line at CH = 32) instead of CH + 4
    s.xc = s.use_lds || nchunk > 1 || tuning_get("dd_fp_extra_channel", 0) == 1;
    // measured on MI355X (8 to 32 slices of 256x256 or 512x512): one float4
    // of padding per pixel from 16 channels on (a 128-byte pixel stride puts
    // every lane's loads on the same offsets within cache lines)
    s.cs = s.xc ? ((nchunk * CH + 1 + 3) / 4) * 4 : (CH >= 16 ? CH + 4 : CH);
    const int forced_cs = tuning_get("dd_fp_pixel_stride", 0);
    if (!s.xc && forced_cs >= CH && forced_cs % (CH >= 4 ? 4 : 1) == 0) s.cs = forced_cs;
    for (auto& v : views) { s.needY |= (v.flags & 1) != 0; s.needX |= (v.flags & 2) != 0; }
    // the branchless parallel FP has no force-FBP reverse weighting
    // (cuda/distanceDriven2DBranchless.cu:321-360)
    s.force = (type & DD_FORCE_FBP) != 0 && !branchless;
    // lanes per column: the split kernel needs one chunk without the next
    // chunk's channel, whole float4s per lane and unpadded CH-float pixels.
    // Default: two float4s per lane (CH / 8 lanes) from 16 channels on, one
    // lane per column below (measured on MI355X, fan FP: 256^2 x 32, 492
    // views: 4 lanes 0.59 ms, 2 lanes 0.78, one lane with two row segments
    // 0.97; 16 channels at 256^2/984 views and 384^2/720: 2 lanes 0.73 and
    // 0.88 ms, 4 lanes 0.89 and 1.12; 8 channels at 256^2/492: one lane with
    // two row segments 0.19 ms, 2 lanes 0.25)
    // row segments of the one-lane kernel: two when one gives it fewer than
    // 4 waves per SIMD (4096 on 1024 SIMDs) and each segment keeps at least
    // 32 rows (measured at 8 slices of 256x256, 492 views, 444 columns: 1
    // segment 0.46 ms, 2 segments 0.25 ms; 4 segments, two per wavefront,
    // 0.25 ms); below 16 channels such grids keep the one-lane kernel even
    // where split lanes are asked for (2 lanes 0.27 ms there)
    const int forced = tuning_get("dd_fp_segments", 0);
    const int rows_min = std::min(g.nx, g.ny);
    const int seg_rule = (rows_min / 2 >= 32 && 2L * ceil_div(d.nu, 128) * nview * nchunk < 4096) ? 1 : 0;
    const int lpc = lpc_req > 0 ? std::min(lpc_req, std::max(CH / 4, 1))
                                : tuning_get("dd_fp_lanes_per_column", CH >= 16 ? CH / 8 : 1);
    const bool split = !s.xc && forced != 2 && (CH >= 16 || seg_rule == 0) && (CH == 8 || CH == 16 || CH == 32) &&
                       (lpc == 2 || lpc == 4 || lpc == 8) && CH % (4 * lpc) == 0;
    int lseg = split ? 0 : seg_rule;
    if (forced == 1 || forced == 2) lseg = forced - 1;
    s.lseg = lseg;
    if (split) {
        s.lpc = lpc;
        s.cs = CH;
    }
    s.rsY = (long)g.nx * s.cs;
    s.rsX = (long)g.ny * s.cs;
    return s;
}

struct DDFPDev {
    const DDView* views;
    const FPRow* rows;
    const int* start;
    const FPGroup* groups;
};

// The forward projector's launches (no host work): imgY and imgX are scratch
// of npix * cs floats each.
static void dd_fp_run(const DDFPSetup& s, const DDFPDev& t, float* prj, const float* img, float* imgY, float* imgX,
                      const Grid& g, const Det& d, int nview, float dsd, float dso, bool fan, hipStream_t stream) {
    const int C = s.C, CH = s.CH, nchunk = s.nchunk, cs = s.cs;
    const DDView* dviews = t.views;
    const FPRow* drows = t.rows;
    const int* dstart = t.start;
    const FPGroup* dgroups = t.groups;
    const bool force = s.force;
    dim3 tb(256);
    if (s.needY) {
        TimedLaunch tt("dd_transpose", stream);
        to_channel_last<<<dim3(ceil_div(g.nx, 32), ceil_div(cs, 32), g.ny), tb, 0, stream>>>(imgY, img, C, cs, g.nx, g.ny, false, (size_t)s.rsY);
        CTP_LAUNCH_CHECK();
    }
    if (s.needX) {
        TimedLaunch tt("dd_transpose", stream);
        to_channel_last<<<dim3(ceil_div(g.nx, 32), ceil_div(cs, 32), g.ny), tb, 0, stream>>>(imgX, img, C, cs, g.nx, g.ny, true, (size_t)s.rsX);
        CTP_LAUNCH_CHECK();
    }
    if (s.use_lds) {
        const int wmax = s.wmax;
        const int CHP = ((CH + 1) + 3) / 4 * 4;
        dim3 grid(ceil_div(d.nu, 64), (unsigned)s.groups.size(), nchunk), block(64 * kFPViews);
        const size_t smem = 4 * std::max<size_t>((size_t)2 * wmax * CHP, (size_t)64 * kFPViews * (CH + 2));
        TimedLaunch tt(fan ? "dd_fan_fp" : "dd_par_fp", stream);
#define CTP_FPL(N)                                                                                              \
    if (CH == N) {                                                                                             \
        if (fan) dd_fp_lds_kernel<N, true><<<grid, block, smem, stream>>>(prj, imgY, imgX, dviews, dgroups,      \
                                                                    drows, dstart, cs, s.rsY, s.rsX, wmax, g, d, nview, dsd, \
                                                                    dso, force);                                 \
        else dd_fp_lds_kernel<N, false><<<grid, block, smem, stream>>>(prj, imgY, imgX, dviews, dgroups,         \
                                                                     drows, dstart, cs, s.rsY, s.rsX, wmax, g, d, nview, dsd, \
                                                                     dso, force);                                \
    }
        CTP_FPL(1) CTP_FPL(2) CTP_FPL(4) CTP_FPL(8) CTP_FPL(16) CTP_FPL(32)
#undef CTP_FPL
        CTP_LAUNCH_CHECK();
        return;
    }
    if (s.lpc > 1) {
        const int nc = 64 / s.lpc;
        dim3 grid(ceil_div(d.nu, nc), ceil_div(nview, 4)), block(256);
        const int nrow = (int)s.rows.size();
        TimedLaunch tt(fan ? "dd_fan_fp" : "dd_par_fp", stream);
#define CTP_FPS_LAUNCH(N, L)                                                                                   \
    if (CH == N && s.lpc == L) {                                                                               \
        if (fan) dd_fp_split_kernel<N, L, true><<<grid, block, 0, stream>>>(prj, imgY, imgX, dviews, drows, nrow, g, d, \
                                                                        nview, dsd, dso, force);               \
        else dd_fp_split_kernel<N, L, false><<<grid, block, 0, stream>>>(prj, imgY, imgX, dviews, drows, nrow, g, d, \
                                                                         nview, dsd, dso, force);              \
    }
        CTP_FPS_LAUNCH(8, 2) CTP_FPS_LAUNCH(16, 2) CTP_FPS_LAUNCH(16, 4) CTP_FPS_LAUNCH(32, 2) CTP_FPS_LAUNCH(32, 4)
        CTP_FPS_LAUNCH(32, 8)
#undef CTP_FPS_LAUNCH
        CTP_LAUNCH_CHECK();
        return;
    }
    const int lseg = s.lseg;
    const int tv = tuning_get("dd_fp_views_per_block", 4);
    const int vpb = (lseg == 0 && (tv == 2 || tv == 4)) ? tv : 1;
    dim3 grid(ceil_div(d.nu, 128 >> lseg), nview, nchunk), block(128);
    if (vpb > 1) {
        grid = dim3(ceil_div(d.nu, 64), ceil_div(nview, vpb), nchunk);
        block = dim3(64 * vpb);
    }
    TimedLaunch tt(fan ? "dd_fan_fp" : "dd_par_fp", stream);
#define CTP_FP_LAUNCH(N, F, X) \
    if (vpb == 4) dd_fp_kernel<N, F, X, 4><<<grid, block, 0, stream>>>(prj, imgY, imgX, dviews, drows, dstart, cs, g, d, nview, dsd, dso, force, lseg); \
    else if (vpb == 2) dd_fp_kernel<N, F, X, 2><<<grid, block, 0, stream>>>(prj, imgY, imgX, dviews, drows, dstart, cs, g, d, nview, dsd, dso, force, lseg); \
    else dd_fp_kernel<N, F, X, 1><<<grid, block, 0, stream>>>(prj, imgY, imgX, dviews, drows, dstart, cs, g, d, nview, dsd, dso, force, lseg)
#define CTP_FP_CASE(N)                                                                                          \
    case N:                                                                                                     \
        if (fan) { if (s.xc) CTP_FP_LAUNCH(N, true, true); else CTP_FP_LAUNCH(N, true, false); }               \
        else { if (s.xc) CTP_FP_LAUNCH(N, false, true); else CTP_FP_LAUNCH(N, false, false); }                 \
        break;
    switch (CH) {
        CTP_FP_CASE(1) CTP_FP_CASE(2) CTP_FP_CASE(4) CTP_FP_CASE(8) CTP_FP_CASE(16) CTP_FP_CASE(32)
        default: fail("dd fp: bad channel chunk");
    }
#undef CTP_FP_CASE
#undef CTP_FP_LAUNCH
    CTP_LAUNCH_CHECK();
}

static std::vector<DDView> dd_views_for(const float* deg_host, int nview, const Grid& g, int type) {
    std::vector<DDView> views = make_dd_views(deg_host, nview, g);
    if ((type & DD_BRANCHLESS) != 0) hard_switch(views);
    return views;
}

// Scratch of one forward projection call: the two channel-last image copies
// and the view / row / chunk / group tables, carved in this order from one
// 256-byte aligned block (each piece rounded up to 256 bytes).
static size_t dd_fp_scratch_bytes(const DDFPSetup& s, const Grid& g, int nview) {
    Plan pl;
    pl.add<float>((size_t)g.ny * s.rsY);
    pl.add<float>((size_t)g.nx * s.rsX);
    pl.add<DDView>(nview);
    pl.add<FPRow>(s.rows.size());
    pl.add<int>(s.nchunk + 1);
    pl.add<FPGroup>(s.groups.size());
    return pl.bytes;
}

// Carves the scratch block, uploads the tables and enqueues the launches; the
// caller synchronises 'stream' before the host tables in 's' and 'views' go.
static void dd_fp_with_scratch(char* base, const DDFPSetup& s, const std::vector<DDView>& views, float* prj,
                               const float* img, const Grid& g, const Det& d, int nview, float dsd, float dso,
                               bool fan, hipStream_t stream) {
    auto take = [&base](size_t bytes) {
        char* p = base;
        base += ((bytes + 255) / 256) * 256;
        return static_cast<void*>(p);
    };
    float* imgY = static_cast<float*>(take(sizeof(float) * (size_t)g.ny * s.rsY));
    float* imgX = static_cast<float*>(take(sizeof(float) * (size_t)g.nx * s.rsX));
    DDView* dviews = static_cast<DDView*>(take(sizeof(DDView) * nview));
    FPRow* drows = static_cast<FPRow*>(take(sizeof(FPRow) * s.rows.size()));
    int* dstart = static_cast<int*>(take(sizeof(int) * (s.nchunk + 1)));
    FPGroup* dgroups = static_cast<FPGroup*>(take(sizeof(FPGroup) * s.groups.size()));
    CTP_CHECK(hipMemcpyAsync(dviews, views.data(), sizeof(DDView) * nview, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(drows, s.rows.data(), sizeof(FPRow) * s.rows.size(), hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dstart, s.start.data(), sizeof(int) * (s.nchunk + 1), hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dgroups, s.groups.data(), sizeof(FPGroup) * s.groups.size(), hipMemcpyHostToDevice,
                             stream));
    dd_fp_run(s, DDFPDev{dviews, drows, dstart, dgroups}, prj, img, imgY, imgX, g, d, nview, dsd, dso, fan, stream);
}


// ---------------------------------------------------------------------------
// Allocation-free forward projection
// ---------------------------------------------------------------------------
// The same launches as dd2d_fp with their scratch in a file-scope device array
// instead of the workspace arena, so the call allocates no device memory.
// Every GPU the library runs on holds its own copy of the array until the
// process ends; calls on one GPU share it (each call overwrites it completely
// before reading: the channel-last copies write every channel, the tables are
// uploaded), so calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[kDDFPNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mutex;
std::map<int, char*> g_scratch_ptr;   // device -> address of 'scratch' on that device
}

// Address of 'scratch' on the current device, looked up once per device (the
// array's name is a host-side stand-in, not a device address).
static char* noalloc_scratch() {
    int dev = 0;
    CTP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* ptr = nullptr;
    CTP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) fail("no device address for the allocation-free scratch array");
    g_scratch_ptr[dev] = static_cast<char*>(ptr);
    return static_cast<char*>(ptr);
}

void dd2d_fp_noalloc(float* prj, const float* img, const float* deg_host, int B, const Grid& g, const Det& d,
                     int nview, float dsd, float dso, int type, bool fan, hipStream_t stream) {
    const std::vector<DDView> views = dd_views_for(deg_host, nview, g, type);
    const DDFPSetup s = dd_fp_setup(views, B, g, d, nview, dsd, dso, type, fan);
    const size_t bytes = dd_fp_scratch_bytes(s, g, nview);
    if (bytes > kDDFPNoAllocScratchBytes)
        fail("the allocation-free forward projection needs " + std::to_string(bytes) +
             " bytes of scratch, more than its fixed " + std::to_string(kDDFPNoAllocScratchBytes) +
             " bytes (two channel-last copies of the B*nz slices plus tables); use the allocating entry point");
    char* base = noalloc_scratch();
    TimingSuspended no_events;   // kernel timing would create HIP events
    dd_fp_with_scratch(base, s, views, prj, img, g, d, nview, dsd, dso, fan, stream);
    CTP_CHECK(hipStreamSynchronize(stream));   // host tables above are released on return
}

// ---------------------------------------------------------------------------
// backprojection
// ---------------------------------------------------------------------------
// Weighted projections, one channel per image slice:
//   sino[view][u][c] = w(view, u) * rows(z) of prj[b][view][.][u],  c = b*nz + z
// with w the BP pre-weight (cuda/distanceDriven2DUtils.cu:8-65) or 1 for FBP.
// Box-integral mode: one row, trunc(z/dv + off_v + nv/2 - 0.5), rows outside
// [0, nv) give 0 (cuda/distanceDriven2DBoxInt.cu:350,383).  Branchless mode: the
// clamped linear interpolation between rows of InterpolateXY
// (cuda/distanceDriven2DBranchless.cu:444-447, cuda/cudaMath.cu:227-260); the
// u-interpolation of the prefix sums there is the exact box integral used below.
// Block (bx, by, bz) = (32 columns, 32 channels, view).
__device__ inline void dd_bp_prep_block(float* __restrict__ sino, const float* __restrict__ prj,
                                        const DDView* __restrict__ views, const SliceRow* __restrict__ slice_row, int C,
                                        int cs, int nz, const Det& d, int nview, float dx, float dy, bool fan,
                                        bool preweight, int bx, int by, int bz) {
    __shared__ float tile[32][33];
    const int u0 = bx * 32, c0 = by * 32, view = bz;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const DDView V = preweight ? views[view] : DDView{};
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, u = u0 + tx;
        float val = 0.f;
        if (c < C && u < d.nu) {
            const int b = c / nz, z = c - b * nz;
            const SliceRow R = slice_row[z];
            if (R.r0 >= 0) {
                const float* pb = prj + ((size_t)b * nview + view) * d.nv * d.nu + u;
                val = pb[(size_t)R.r0 * d.nu];
                if (R.w1 != 0.f) val = (1.f - R.w1) * val + R.w1 * pb[(size_t)R.r1 * d.nu];
                if (preweight) {
                    const float a = fan ? (-((float)u - (float)(d.nu - 1) / 2.0f) - d.off_u) * d.du : 0.f;
                    val *= (V.flags & 4) ? dy / fabsf(cosf(V.deg - a)) : dx / fabsf(sinf(V.deg - a));
                }
            }
        }
        tile[k][tx] = val;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int u = u0 + k, c = c0 + tx;
        if (u < d.nu && c < cs) sino[((size_t)view * d.nu + u) * cs + c] = tile[tx][k];
    }
}

__global__ __launch_bounds__(256) void dd_bp_prep_kernel(float* __restrict__ sino, const float* __restrict__ prj,
                                                         const DDView* __restrict__ views, const SliceRow* __restrict__ slice_row,
                                                         int C, int cs, int nz, Det d, int nview, float dx, float dy,
                                                         bool fan, bool preweight) {
    dd_bp_prep_block(sino, prj, views, slice_row, C, cs, nz, d, nview, dx, dy, fan, preweight, blockIdx.x, blockIdx.y,
                     blockIdx.z);
}

void slices_to_channel_last(float* sino, const float* prj, const SliceRow* rows_dev, int C, int cs, int nz,
                            const Det& d, int nview, hipStream_t stream) {
    TimedLaunch t("fbp_bp_prep", stream);
    dd_bp_prep_kernel<<<dim3(ceil_div(d.nu, 32), ceil_div(cs, 32), nview), 256, 0, stream>>>(
        sino, prj, nullptr, rows_dev, C, cs, nz, d, nview, 1.f, 1.f, false, false);
    CTP_LAUNCH_CHECK();
}

// Detector footprint [a1, a2] (cell units, unclamped) of a pixel applied to the
// channel vectors of one view: IntegralBoxX over a detector row divided by
// (a2 - a1), cuda/distanceDriven2DBoxInt.cu:375-383.
template <int CH>
__device__ inline void dd_cells(float* acc, const float* __restrict__ base, int cs, float a1, float a2, int n, float scale) {
    const float lo = min_f(a1, a2), hi = max_f(a1, a2);
    const float a = med_clamp(lo, 0.f, (float)n), b = med_clamp(hi, 0.f, (float)n);
    if (!(b > a)) return;
    const float inv = fast_div(scale, hi - lo);
    const int i1 = min((int)a, n - 1), i2 = min((int)b, n - 1);
    for (int i = i1; i <= i2; i++) {
        const float w = (fminf(b, (float)(i + 1)) - fmaxf(a, (float)i)) * inv;
        const float* p = base + (size_t)i * cs;
        if constexpr (CH % 4 == 0) {
#pragma unroll
            for (int j = 0; j < CH; j += 4) {
                const float4 q = *reinterpret_cast<const float4*>(p + j);
                acc[j] = fmaf(w, q.x, acc[j]);
                acc[j + 1] = fmaf(w, q.y, acc[j + 1]);
                acc[j + 2] = fmaf(w, q.z, acc[j + 2]);
                acc[j + 3] = fmaf(w, q.w, acc[j + 3]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) acc[j] = fmaf(w, p[j], acc[j]);
        }
    }
}

// Block (bx, by, bz) = (16 x 16 pixels, view group x channel chunk).
template <int CH, bool FAN>
__device__ inline void dd_bp_block(float* __restrict__ img, float* __restrict__ part, const float* __restrict__ sino,
                                   const DDView* __restrict__ views, int C, int cs, int nchunk, int vpg, const Grid& g,
                                   const Det& d, int nview, float dso, bool fbp, int bx, int by, int bz) {
    const int ix = bx * 16 + (threadIdx.x & 15);
    const int iy = by * 16 + (threadIdx.x >> 4);
    const int grp = bz / nchunk;
    const int c0 = (bz - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;
    // pixel centre and edges, cuda/distanceDriven2DBoxInt.cu:340-349
    const float xc = ((float)ix - g.nx / 2.f) * g.dx + g.cx + g.dx / 2;
    const float yc = ((float)iy - g.ny / 2.f) * g.dy + g.cy + g.dy / 2;
    const float x1 = xc - g.dx / 2.0f, x2 = xc + g.dx / 2.0f;
    const float y1 = yc - g.dy / 2.0f, y2 = yc + g.dy / 2.0f;
    float acc[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) acc[j] = 0.f;
    const float half = (float)d.nu / 2.0f;
    const size_t vstride = (size_t)d.nu * cs;
    for (int view = vbeg; view < vend; view++) {
        const DDView V = views[view];
        const float c = V.c, s = V.s;
        const float* __restrict__ base = sino + view * vstride + c0;
        if (FAN) {
            // cuda/distanceDriven2DBoxInt.cu:362-391 and GetProjectionOnDetector (Utils.cu:96-102)
            float p1x, p1y, p2x, p2y;
            if (V.flags & 4) { p1x = x1; p1y = yc; p2x = x2; p2y = yc; }
            else { p1x = xc; p1y = y1; p2x = xc; p2y = y2; }
            const float a1 = atanf(fast_div(p1x * c + p1y * s, -p1x * s + p1y * c + dso));
            const float a2 = atanf(fast_div(p2x * c + p2y * s, -p2x * s + p2y * c + dso));
            const float u1 = -(fast_div(a1, d.du) + d.off_u) + half;
            const float u2 = -(fast_div(a2, d.du) + d.off_u) + half;
            float scale = 1.f;
            if (fbp) {
                const float rx = xc * c + yc * s, ry = -xc * s + yc * c;
                scale = fast_div(dso * dso, rx * rx + (dso + ry) * (dso + ry));
            }
            dd_cells<CH>(acc, base, cs, u1, u2, d.nu, scale);
        } else {
            // cuda/distanceDriven2DBoxInt.cu:439-493 (blended main axes)
            if (V.flags & 1) {
                const float u1 = -(fast_div(x1 * c + yc * s, d.du) + d.off_u) + half;
                const float u2 = -(fast_div(x2 * c + yc * s, d.du) + d.off_u) + half;
                dd_cells<CH>(acc, base, cs, u1, u2, d.nu, V.wx);
            }
            if (V.flags & 2) {
                const float u1 = -(fast_div(xc * c + y1 * s, d.du) + d.off_u) + half;
                const float u2 = -(fast_div(xc * c + y2 * s, d.du) + d.off_u) + half;
                dd_cells<CH>(acc, base, cs, u1, u2, d.nu, 1.f - V.wx);
            }
        }
    }
    store_bp_result<CH>(img, part, acc, grp, c0, C, (size_t)g.nx * g.ny, (size_t)iy * g.nx + ix);
}

template <int CH, bool FAN>
__global__ __launch_bounds__(256) void dd_bp_kernel(float* __restrict__ img, float* __restrict__ part,
                                                    const float* __restrict__ sino, const DDView* __restrict__ views,
                                                    int C, int cs, int nchunk, int vpg, Grid g, Det d, int nview,
                                                    float dso, bool fbp) {
    dd_bp_block<CH, FAN>(img, part, sino, views, C, cs, nchunk, vpg, g, d, nview, dso, fbp, blockIdx.x, blockIdx.y,
                         blockIdx.z);
}

// The backprojection above with LPC lanes per pixel, for the reason given at
// dd_fp_split_kernel: lane s of a pixel holds the float4s at channel offsets
// 4 (s + LPC k), so a load instruction reads one contiguous 16 LPC-byte piece
// of each detector cell, and a wavefront's 64 / LPC pixels (a WX x WY tile)
// cover fewer cells.  In the fan beam the views' footprints (two atans each)
// are spread over the pixel's lanes.  A block is 2 x 2 wavefront tiles.  Same arithmetic per channel as
// dd_bp_kernel.
// Block (bx, by, bz) = (2 x 2 wavefront tiles of pixels, view group x channel chunk).
template <int CH, int LPC, bool FAN>
__device__ inline void dd_bp_split_block(float* __restrict__ img, float* __restrict__ part,
                                         const float* __restrict__ sino, const DDView* __restrict__ views, int C,
                                         int cs, int nchunk, int vpg, const Grid& g, const Det& d, int nview, float dso,
                                         bool fbp, int bx, int by, int bz) {
    static_assert(CH % (4 * LPC) == 0 && LPC >= 2, "a lane holds whole float4s");
    constexpr int NQ = CH / (4 * LPC);
    constexpr int WX = LPC == 2 ? 8 : 4, WY = 64 / LPC / WX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = lane / LPC, sub = lane % LPC;
    const int ix = bx * (2 * WX) + (wave & 1) * WX + p % WX;
    const int iy = by * (2 * WY) + (wave >> 1) * WY + p / WX;
    const int grp = bz / nchunk;
    const int c0 = (bz - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;   // the lanes of a pixel leave together
    // pixel centre and edges, cuda/distanceDriven2DBoxInt.cu:340-349
    const float xc = ((float)ix - g.nx / 2.f) * g.dx + g.cx + g.dx / 2;
    const float yc = ((float)iy - g.ny / 2.f) * g.dy + g.cy + g.dy / 2;
    const float x1 = xc - g.dx / 2.0f, x2 = xc + g.dx / 2.0f;
    const float y1 = yc - g.dy / 2.0f, y2 = yc + g.dy / 2.0f;
    float acc[4 * NQ];
#pragma unroll
    for (int j = 0; j < 4 * NQ; j++) acc[j] = 0.f;
    const float half = (float)d.nu / 2.0f;
    const size_t vstride = (size_t)d.nu * cs;
    if (FAN) {
        // lane s of a pixel evaluates the footprint of view v0 + s; the pixel's
        // lanes read each view's footprint from it
        const int lead = lane & ~(LPC - 1);
        for (int v0 = vbeg; v0 < vend; v0 += LPC) {
            float mu1 = 0.f, mu2 = 0.f, msc = 1.f;
            if (v0 + sub < vend) {
                // cuda/distanceDriven2DBoxInt.cu:362-391 and GetProjectionOnDetector (Utils.cu:96-102)
                const DDView V = views[v0 + sub];
                const float c = V.c, s = V.s;
                float p1x, p1y, p2x, p2y;
                if (V.flags & 4) { p1x = x1; p1y = yc; p2x = x2; p2y = yc; }
                else { p1x = xc; p1y = y1; p2x = xc; p2y = y2; }
                const float a1 = atanf(fast_div(p1x * c + p1y * s, -p1x * s + p1y * c + dso));
                const float a2 = atanf(fast_div(p2x * c + p2y * s, -p2x * s + p2y * c + dso));
                mu1 = -(fast_div(a1, d.du) + d.off_u) + half;
                mu2 = -(fast_div(a2, d.du) + d.off_u) + half;
                if (fbp) {
                    const float rx = xc * c + yc * s, ry = -xc * s + yc * c;
                    msc = fast_div(dso * dso, rx * rx + (dso + ry) * (dso + ry));
                }
            }
#pragma unroll
            for (int t = 0; t < LPC; t++) {
                const int view = v0 + t;
                if (view >= vend) break;   // uniform
                // broadcast from lane t of the pixel's lanes (DPP within the quad)
                const float u1 = group_bcast<LPC>(mu1, t, lead), u2 = group_bcast<LPC>(mu2, t, lead);
                const float scale = fbp ? group_bcast<LPC>(msc, t, lead) : 1.f;
                dd_line_split<NQ, LPC>(acc, sino + view * vstride + c0, 4 * sub, cs, u1, u2, d.nu, scale);
            }
        }
    } else {
        for (int view = vbeg; view < vend; view++) {
            const DDView V = views[view];
            const float c = V.c, s = V.s;
            const float* __restrict__ base = sino + view * vstride + c0;
            // cuda/distanceDriven2DBoxInt.cu:439-493 (blended main axes)
            if (V.flags & 1) {
                const float u1 = -(fast_div(x1 * c + yc * s, d.du) + d.off_u) + half;
                const float u2 = -(fast_div(x2 * c + yc * s, d.du) + d.off_u) + half;
                dd_line_split<NQ, LPC>(acc, base, 4 * sub, cs, u1, u2, d.nu, V.wx);
            }
            if (V.flags & 2) {
                const float u1 = -(fast_div(xc * c + y1 * s, d.du) + d.off_u) + half;
                const float u2 = -(fast_div(xc * c + y2 * s, d.du) + d.off_u) + half;
                dd_line_split<NQ, LPC>(acc, base, 4 * sub, cs, u1, u2, d.nu, 1.f - V.wx);
            }
        }
    }
    const size_t npix = (size_t)g.nx * g.ny, pix = (size_t)iy * g.nx + ix;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int cc = c0 + 4 * (sub + LPC * k) + e;
            if (cc >= C) continue;
            if (part) part[((size_t)grp * C + cc) * npix + pix] = acc[4 * k + e];
            else img[(size_t)cc * npix + pix] += acc[4 * k + e];
        }
    }
}

template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void dd_bp_split_kernel(float* __restrict__ img, float* __restrict__ part,
                                                          const float* __restrict__ sino,
                                                          const DDView* __restrict__ views, int C, int cs, int nchunk,
                                                          int vpg, Grid g, Det d, int nview, float dso, bool fbp) {
    dd_bp_split_block<CH, LPC, FAN>(img, part, sino, views, C, cs, nchunk, vpg, g, d, nview, dso, fbp, blockIdx.x,
                                    blockIdx.y, blockIdx.z);
}

// Number of view groups a pixel-driven backprojection is split into so that
// the launch has enough waves in flight to hide the per-view load latency
// (one lane per pixel alone gives 4 waves per SIMD at 512x512).  Partial
// images of the groups are summed in a fixed order afterwards: deterministic.
// Lanes per pixel of the DD backprojection for a chunk of CH channels (the
// dd_bp_lanes_per_pixel key, else two float4 per lane: 4 lanes at 32
// channels, 2 at 16 and at 8).  Measured on MI355X, fan 984 views: at
// 512x512x32 one lane 5.09 ms, two 3.19, four 3.07; at 16 channels two lanes
// against four 0.45 / 0.55 ms (256^2), 1.48 / 1.96 ms (512^2), and at 384^2
// with 720 views 0.58 / 0.72 ms.
static int dd_bp_lanes(int CH) {
    const int t = tuning_get("dd_bp_lanes_per_pixel", 0);
    if (t > 0) return t == 4 && CH % 16 != 0 ? 2 : t;
    return CH >= 32 ? 4 : 2;
}

int bp_view_groups(size_t npix, int nchunk, int nview) {
    // The pixel-driven backprojections launch one 256-lane block per 16x16
    // pixels and channel chunk.  Below 1024 blocks (4 per CU) the launch
    // cannot hide its gather latency, so the views are split into groups
    // (enough for 2048 blocks) whose partial images are summed afterwards.
    // Measured on MI355X: at 256x256x8 groups of 8 beat 4 and 16 for the
    // adjoint; at 512x512x32 (1024 blocks) the partial images cost more than
    // extra waves gain, so such launches keep one group.
    const size_t blocks = std::max<size_t>(1, (npix + 255) / 256 * (size_t)std::max(nchunk, 1));
    const int automatic = blocks >= 1024 ? 1 : (int)std::min<size_t>(64, (2048 + blocks - 1) / blocks);
    int k = tuning_get("bp_view_groups", automatic);
    k = std::max(1, std::min(k, std::max(1, nview / 8)));
    k = std::max(1, std::min(k, 65535 / std::max(nchunk, 1)));   // grid.z = nchunk * groups
    return k;
}

__global__ __launch_bounds__(256) void reduce_parts_kernel(float* __restrict__ img, const float* __restrict__ part,
                                                           int groups, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < groups; k++) s += part[(size_t)k * n + i];
    img[i] += s;
}

void reduce_parts(float* img, const float* part, int groups, size_t n, hipStream_t stream) {
    reduce_parts_kernel<<<(unsigned)ceil_div((long)n, 256), 256, 0, stream>>>(img, part, groups, n);
    CTP_LAUNCH_CHECK();
}

struct DDBPSetup {
    int C = 0, CH = 1, nchunk = 1, cs = 0, groups = 1, vpg = 1;
    bool fbp = false;
    std::vector<SliceRow> srow;
};

static DDBPSetup dd_bp_setup(int B, const Grid& g, const Det& d, int nview, int type) {
    const bool branchless = (type & DD_BRANCHLESS) != 0;
    DDBPSetup s;
    // the branchless BP ignores the force-FBP bit (cuda/distanceDriven2DBranchless.cu:547)
    s.fbp = branchless ? (type & DD_FBP) != 0 : (type & (DD_FBP | DD_FORCE_FBP)) != 0;
    s.C = B * g.nz;
    s.CH = pick_ch(s.C);
    s.nchunk = ceil_div(s.C, s.CH);
    s.cs = s.nchunk * s.CH;
    // detector rows of each slice, cuda/distanceDriven2DBoxInt.cu:340-350
    s.srow.resize(g.nz);
    for (int z = 0; z < g.nz; z++) {
        const double zc = (z - g.nz / 2.0) * g.dz + g.cz + g.dz / 2.0;
        const double ivf = zc / d.dv + d.off_v + d.nv / 2.0 - 0.5;
        SliceRow r{-1, -1, 0.f, 0};
        if (!branchless) {
            const int t = (int)ivf;   // the float row is passed to an int parameter: truncation
            if (t >= 0 && t < d.nv) r = SliceRow{t, t, 0.f, 0};
        } else {
            const double y = std::min((double)d.nv, std::max(0.0, ivf));
            const int iy = (int)y;
            const int r0 = clampi(iy, 0, d.nv), r1 = clampi(iy + 1, 0, d.nv);
            r = SliceRow{r0, r1, (float)(y - r0), 0};
        }
        s.srow[z] = r;
    }
    s.groups = bp_view_groups((size_t)g.nx * g.ny, s.nchunk, nview);
    s.vpg = ceil_div(nview, s.groups);
    return s;
}

// The backprojector's launches: sino is scratch of nview * nu * cs floats,
// part of groups * C * npix floats when groups > 1 (else null).
static void dd_bp_run(const DDBPSetup& s, const DDView* dviews, const SliceRow* dsrow, float* sino, float* part,
                      float* img, const float* prj, const Grid& g, const Det& d, int nview, float dso, bool fan,
                      hipStream_t stream) {
    const int C = s.C, CH = s.CH, nchunk = s.nchunk, cs = s.cs, groups = s.groups, vpg = s.vpg;
    const bool fbp = s.fbp;
    {
        TimedLaunch t("dd_bp_prep", stream);
        dd_bp_prep_kernel<<<dim3(ceil_div(d.nu, 32), ceil_div(cs, 32), nview), 256, 0, stream>>>(
            sino, prj, dviews, dsrow, C, cs, g.nz, d, nview, g.dx, g.dy, fan, !fbp);
        CTP_LAUNCH_CHECK();
    }
    TimedLaunch tb(fan ? "dd_fan_bp" : "dd_par_bp", stream);
    int lpc = dd_bp_lanes(CH);
    if ((lpc == 2 || lpc == 4) && CH >= 8 && CH % (4 * lpc) == 0) {
        const int bx = lpc == 2 ? 16 : 8, by = lpc == 2 ? 8 : 8;
        dim3 grid(ceil_div(g.nx, bx), ceil_div(g.ny, by), nchunk * groups), block(256);
#define CTP_BPS_LAUNCH(N, L)                                                                                  \
    if (CH == N && lpc == L) {                                                                                \
        if (fan) dd_bp_split_kernel<N, L, true><<<grid, block, 0, stream>>>(img, part, sino, dviews, C, cs, nchunk, \
                                                                        vpg, g, d, nview, dso, fbp);           \
        else dd_bp_split_kernel<N, L, false><<<grid, block, 0, stream>>>(img, part, sino, dviews, C, cs, nchunk, \
                                                                         vpg, g, d, nview, dso, fbp);          \
    }
        CTP_BPS_LAUNCH(8, 2) CTP_BPS_LAUNCH(16, 2) CTP_BPS_LAUNCH(16, 4) CTP_BPS_LAUNCH(32, 2) CTP_BPS_LAUNCH(32, 4)
#undef CTP_BPS_LAUNCH
        CTP_LAUNCH_CHECK();
        if (part) reduce_parts(img, part, groups, (size_t)C * g.nx * g.ny, stream);
        return;
    }
    dim3 grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), nchunk * groups), block(256);
#define CTP_BP_CASE(N)                                                                                        \
    case N:                                                                                                    \
        if (fan) dd_bp_kernel<N, true><<<grid, block, 0, stream>>>(img, part, sino, dviews, C, cs, nchunk, vpg, g, d, \
                                                                   nview, dso, fbp);                                \
        else dd_bp_kernel<N, false><<<grid, block, 0, stream>>>(img, part, sino, dviews, C, cs, nchunk, vpg, g, d,  \
                                                                nview, dso, fbp);                                   \
        break;
    switch (CH) {
        CTP_BP_CASE(1) CTP_BP_CASE(2) CTP_BP_CASE(4) CTP_BP_CASE(8) CTP_BP_CASE(16) CTP_BP_CASE(32)
        default: fail("dd bp: bad channel chunk");
    }
#undef CTP_BP_CASE
    CTP_LAUNCH_CHECK();
    if (part) reduce_parts(img, part, groups, (size_t)C * g.nx * g.ny, stream);
}

// ---------------------------------------------------------------------------
// geometry tables of the per-call entry points, cached per device
// ---------------------------------------------------------------------------
// The reference rebuilds its tables on every call; iterative reconstructions
// call the same geometry and angles many times.  The view table (with its FP64
// axis decisions), the forward projector's row / chunk / group tables and the
// backprojector's slice rows of one (geometry, batch, angles, typeProjector,
// tuning generation) are kept on the device and reused; the per-call scratch
// (channel-last copies, weighted projections, partial images) still comes
// from the stream's workspace.  Entries are pinned while a call uses them and
// the least recently used unpinned ones are freed beyond kDDCallCache.
constexpr size_t kDDCallCache = 16;

struct DDCallTables {
    int device = 0, B = 0, nview = 0, type = 0;
    bool fan = true;
    Grid g{};
    Det d{};
    float dsd = 0.f, dso = 0.f;
    long gen = 0;
    std::vector<float> deg;
    DDFPSetup fp;
    DDBPSetup bp;
    void* mem = nullptr;
    DDView* views = nullptr;
    FPRow* rows = nullptr;
    int* start = nullptr;
    FPGroup* groups = nullptr;
    SliceRow* srow = nullptr;
    int pins = 0;
    unsigned long long used = 0;
    bool matches(int dev, const float* a, int B_, const Grid& g_, const Det& d_, int nv, float dsd_, float dso_,
                 int type_, bool fan_, long gen_) const {
        return dev == device && B_ == B && nv == nview && type_ == type && fan_ == fan && gen_ == gen &&
               std::memcmp(&g_, &g, sizeof(Grid)) == 0 && std::memcmp(&d_, &d, sizeof(Det)) == 0 &&
               std::memcmp(&dsd_, &dsd, sizeof(float)) == 0 && std::memcmp(&dso_, &dso, sizeof(float)) == 0 &&
               std::memcmp(a, deg.data(), sizeof(float) * nv) == 0;
    }
};

namespace {
std::mutex g_ddc_mutex;
std::vector<std::unique_ptr<DDCallTables>> g_ddc;
unsigned long long g_ddc_clock = 0;
}

// The entry for this call, built on a miss; pinned until dd_tables_release.
static DDCallTables* dd_tables_acquire(const float* deg_host, int B, const Grid& g, const Det& d, int nview,
                                       float dsd, float dso, int type, bool fan) {
    int dev = 0;
    CTP_CHECK(hipGetDevice(&dev));
    const long gen = tuning_generation();
    std::lock_guard<std::mutex> lk(g_ddc_mutex);
    for (auto& e : g_ddc)
        if (e->matches(dev, deg_host, B, g, d, nview, dsd, dso, type, fan, gen)) {
            e->pins++;
            e->used = ++g_ddc_clock;
            return e.get();
        }
    // free the least recently used unpinned entries of this device beyond the
    // limit (entries of other devices and entries in use stay: the cache can
    // then hold more than kDDCallCache entries for a while)
    while (g_ddc.size() >= kDDCallCache) {
        auto victim = g_ddc.end();
        for (auto it = g_ddc.begin(); it != g_ddc.end(); ++it)
            if ((*it)->pins == 0 && (*it)->device == dev && (victim == g_ddc.end() || (*it)->used < (*victim)->used))
                victim = it;
        if (victim == g_ddc.end()) break;
        (void)hipFree((*victim)->mem);
        g_ddc.erase(victim);
    }
    std::unique_ptr<DDCallTables> e(new DDCallTables());
    e->device = dev; e->B = B; e->nview = nview; e->type = type; e->fan = fan; e->g = g; e->d = d;
    e->dsd = dsd; e->dso = dso; e->gen = gen;
    e->deg.assign(deg_host, deg_host + nview);
    const std::vector<DDView> views = dd_views_for(deg_host, nview, g, type);
    e->fp = dd_fp_setup(views, B, g, d, nview, dsd, dso, type, fan);
    e->bp = dd_bp_setup(B, g, d, nview, type);
    Plan pl;
    pl.add<DDView>(nview);
    pl.add<FPRow>(std::max<size_t>(e->fp.rows.size(), 1));
    pl.add<int>(e->fp.nchunk + 1);
    pl.add<FPGroup>(std::max<size_t>(e->fp.groups.size(), 1));
    pl.add<SliceRow>(g.nz);
    CTP_CHECK(hipMalloc(&e->mem, pl.bytes));
    try {
        char* m = static_cast<char*>(e->mem);
        auto carve = [&m](size_t bytes) { void* r = m; m += (bytes + 255) / 256 * 256; return r; };
        e->views = static_cast<DDView*>(carve(sizeof(DDView) * nview));
        e->rows = static_cast<FPRow*>(carve(sizeof(FPRow) * std::max<size_t>(e->fp.rows.size(), 1)));
        e->start = static_cast<int*>(carve(sizeof(int) * (e->fp.nchunk + 1)));
        e->groups = static_cast<FPGroup*>(carve(sizeof(FPGroup) * std::max<size_t>(e->fp.groups.size(), 1)));
        e->srow = static_cast<SliceRow*>(carve(sizeof(SliceRow) * g.nz));
        CTP_CHECK(hipMemcpy(e->views, views.data(), sizeof(DDView) * nview, hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(e->rows, e->fp.rows.data(), sizeof(FPRow) * e->fp.rows.size(), hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(e->start, e->fp.start.data(), sizeof(int) * (e->fp.nchunk + 1), hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(e->groups, e->fp.groups.data(), sizeof(FPGroup) * e->fp.groups.size(),
                            hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(e->srow, e->bp.srow.data(), sizeof(SliceRow) * g.nz, hipMemcpyHostToDevice));
    } catch (...) {
        (void)hipFree(e->mem);   // the entry is not kept
        throw;
    }
    e->pins = 1;
    e->used = ++g_ddc_clock;
    g_ddc.push_back(std::move(e));
    return g_ddc.back().get();
}

// Frees every entry no call is using (on every device).
void release_dd_call_cache() {
    std::lock_guard<std::mutex> lk(g_ddc_mutex);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto it = g_ddc.begin(); it != g_ddc.end();) {
        if ((*it)->pins == 0) {
            (void)hipSetDevice((*it)->device);
            (void)hipFree((*it)->mem);
            it = g_ddc.erase(it);
        } else {
            ++it;
        }
    }
    (void)hipSetDevice(cur);
}

// Unpins the entry when the call's work has completed (or failed).
struct DDTablesPin {
    DDCallTables* t;
    ~DDTablesPin() {
        std::lock_guard<std::mutex> lk(g_ddc_mutex);
        t->pins--;
    }
};

void dd2d_fp(float* prj, const float* img, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
             float dsd, float dso, int type, bool fan, hipStream_t stream) {
    DDTablesPin pin{dd_tables_acquire(deg_host, B, g, d, nview, dsd, dso, type, fan)};
    const DDFPSetup& s = pin.t->fp;
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>((size_t)g.ny * s.rsY);
    pl.add<float>((size_t)g.nx * s.rsX);
    ws.reserve(pl.bytes);
    float* imgY = ws.take_n<float>((size_t)g.ny * s.rsY);
    float* imgX = ws.take_n<float>((size_t)g.nx * s.rsX);
    dd_fp_run(s, DDFPDev{pin.t->views, pin.t->rows, pin.t->start, pin.t->groups}, prj, img, imgY, imgX, g, d, nview,
              dsd, dso, fan, stream);
    CTP_CHECK(hipStreamSynchronize(stream));   // the tables stay pinned until the work is done
}

void dd2d_bp(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
             float dsd, float dso, int type, bool fan, hipStream_t stream) {
    DDTablesPin pin{dd_tables_acquire(deg_host, B, g, d, nview, dsd, dso, type, fan)};
    const DDBPSetup& s = pin.t->bp;
    const size_t npix = (size_t)g.nx * g.ny;
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>((size_t)nview * d.nu * s.cs);
    if (s.groups > 1) pl.add<float>((size_t)s.groups * s.C * npix);
    ws.reserve(pl.bytes);
    float* sino = ws.take_n<float>((size_t)nview * d.nu * s.cs);
    float* part = s.groups > 1 ? ws.take_n<float>((size_t)s.groups * s.C * npix) : nullptr;
    dd_bp_run(s, pin.t->views, pin.t->srow, sino, part, img, prj, g, d, nview, dso, fan, stream);
    CTP_CHECK(hipStreamSynchronize(stream));   // the tables stay pinned until the work is done
}

// ---------------------------------------------------------------------------
// exact adjoint of the forward projector (no counterpart in the reference)
// ---------------------------------------------------------------------------
// img += A^T y for the forward projector above (BoxInt or branchless, fan or
// parallel, blend band, z interpolation, force-FBP reverse weight), for
// gradients.  The reference's backprojector is not the transpose of its
// forward projector (different footprints and weights), so it cannot serve as
// the exact gradient.  Deterministic gather: one lane per pixel, CH slice
// channels; per view and trace the pixel's edge interval is projected onto the
// detector to find the candidate columns, and each candidate's weight is
// recomputed from the same FP32 line parameters the forward projector uses.
// Candidates whose weight is not positive add 0 times their channels, so a
// non-finite detector value also reaches the pixels next to its footprint.
struct FPTrace {           // one trace of a (view, column): edge lines r * (t - o) + b and the norm
    float r1, o1, b1, r2, o2, b2, norm, pad;
};
struct FPLine {            // per (view, column): trace along y (x of row t), trace along x (y of column t)
    FPTrace tr[2];         // 32-byte aligned records: a candidate's trace is two 16-byte loads
};

template <bool FAN>
__global__ __launch_bounds__(256) void dd_fp_lines_kernel(FPLine* __restrict__ lines, const DDView* __restrict__ views,
                                                          Grid g, Det d, int nview, float dsd, float dso) {
    const int iu = blockIdx.x * 256 + threadIdx.x, view = blockIdx.y;
    if (iu >= d.nu) return;
    const DDView V = views[view];
    float2 s1, s2, d1, d2;
    const float a = fp_cell_edges<FAN>(iu, V, g, d, dsd, dso, s1, s2, d1, d2);
    FPLine L{};
    if (V.flags & 1)
        L.tr[0] = FPTrace{(d1.x - s1.x) / (d1.y - s1.y), s1.y, s1.x, (d2.x - s2.x) / (d2.y - s2.y), s2.y, s2.x,
                          V.wx * g.dy / fabsf(cosf(V.deg - a)), 0.f};
    if (V.flags & 2)
        L.tr[1] = FPTrace{(d1.y - s1.y) / (d1.x - s1.x), s1.x, s1.y, (d2.y - s2.y) / (d2.x - s2.x), s2.x, s2.y,
                          (1.f - V.wx) * g.dx / fabsf(sinf(V.deg - a)), 0.f};
    lines[(size_t)view * d.nu + iu] = L;
}

// G[view][u][c] = rev * sum over the detector rows of slice channel c of their
// z-interpolation weight times y (the transpose of the forward projector's
// per-row interpolation of the slice accumulators)
using ChanRow = ChanRowHost;

__global__ __launch_bounds__(256) void dd_adj_rows_kernel(float* __restrict__ G, const float* __restrict__ prj,
                                                          const DDView* __restrict__ views,
                                                          const int* __restrict__ cstart, const ChanRow* __restrict__ crow,
                                                          int C, int cs, int nz, Det d, int nview, bool rev) {
    const int u = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int view = blockIdx.z;
    if (u >= d.nu || c >= cs) return;
    float sum = 0.f;
    if (c < C) {
        const int b = c / nz;
        const float* pb = prj + ((size_t)b * nview + view) * d.nv * d.nu + u;
        for (int k = cstart[c]; k < cstart[c + 1]; k++) sum = fmaf(crow[k].w, pb[(size_t)crow[k].iv * d.nu], sum);
        if (rev) sum *= views[view].rev;
    }
    G[((size_t)view * d.nu + u) * cs + c] = sum;
}

void rows_to_channel_last(float* q, const float* prj, const int* cstart, const ChanRowHost* crow, int C, int cs,
                          int nz, const Det& d, int nview, hipStream_t stream) {
    dd_adj_rows_kernel<<<dim3(ceil_div(d.nu, 64), ceil_div(cs, 4), nview), 256, 0, stream>>>(
        q, prj, nullptr, cstart, crow, C, cs, nz, d, nview, false);
    CTP_LAUNCH_CHECK();
}

// detector column coordinate (column iu spans [iu, iu + 1]) of a physical point
template <bool FAN>
__device__ inline float adj_ucoord(float px, float py, float c, float s, const Det& d, float dso) {
    const float half = (float)d.nu / 2.0f;
    if (FAN) return -(atanf((px * c + py * s) / (-px * s + py * c + dso)) / d.du + d.off_u) + half;
    return -((px * c + py * s) / d.du + d.off_u) + half;
}

template <int CH, bool FAN>
__global__ __launch_bounds__(256) void dd_fp_adjoint_kernel(float* __restrict__ img, float* __restrict__ part,
                                                            const float* __restrict__ G, const FPLine* __restrict__ lines,
                                                            const DDView* __restrict__ views, int C, int cs, int nchunk,
                                                            int vpg, Grid g, Det d, int nview, float dso) {
    const int ix = blockIdx.x * 16 + (threadIdx.x & 15);
    const int iy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int grp = blockIdx.z / nchunk;
    const int c0 = (blockIdx.z - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;
    // pixel edges in physical coordinates
    const float xa = ((float)ix - g.nx / 2.f) * g.dx + g.cx, xb = xa + g.dx;
    const float ya = ((float)iy - g.ny / 2.f) * g.dy + g.cy, yb = ya + g.dy;
    const float xc = xa + g.dx / 2, yc = ya + g.dy / 2;
    float acc[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) acc[j] = 0.f;
    for (int view = vbeg; view < vend; view++) {
        const DDView V = views[view];
        const FPLine* __restrict__ Lv = lines + (size_t)view * d.nu;
        const float* __restrict__ Gv = G + (size_t)view * d.nu * cs + c0;
#pragma unroll
        for (int tr = 0; tr < 2; tr++) {
            if (!(V.flags & (1 << tr))) continue;
            // the pixel's interval on its traced row: [ix, ix+1] at y = iy + 0.5 (tr 0)
            // or [iy, iy+1] at x = ix + 0.5 (tr 1), in index units
            const float ua = tr == 0 ? adj_ucoord<FAN>(xa, yc, V.c, V.s, d, dso) : adj_ucoord<FAN>(xc, ya, V.c, V.s, d, dso);
            const float ub = tr == 0 ? adj_ucoord<FAN>(xb, yc, V.c, V.s, d, dso) : adj_ucoord<FAN>(xc, yb, V.c, V.s, d, dso);
            // candidate columns: those whose span [k, k+1] meets [ua, ub],
            // widened by 1e-3 column for the difference between this FP32
            // projection and the forward projector's FP32 edge lines (about
            // 1e-5 column); each candidate's weight is then exact
            const int k0 = max((int)floorf(fminf(ua, ub) - 1e-3f), 0);
            const int k1 = min((int)floorf(fmaxf(ua, ub) + 1e-3f), d.nu - 1);
            const float t = tr == 0 ? (float)iy + 0.5f : (float)ix + 0.5f;
            const int n = tr == 0 ? g.nx : g.ny;
            const float lo = tr == 0 ? (float)ix : (float)iy;
            for (int k = k0; k <= k1; k++) {
                // the candidate's trace record and its channels load together
                // (one memory round trip per candidate; taking candidates in
                // pairs measured slower); a candidate whose weight is not
                // positive adds 0 times its channels
                const float4* __restrict__ lp = reinterpret_cast<const float4*>(&Lv[k].tr[tr]);
                const float4 l0 = lp[0], l1 = lp[1];   // r1 o1 b1 r2 | o2 b2 norm -
                float q[CH];
                const float* __restrict__ gp = Gv + (size_t)k * cs;
                if constexpr (CH % 4 == 0) {
#pragma unroll
                    for (int j = 0; j < CH; j += 4) {
                        const float4 v = *reinterpret_cast<const float4*>(gp + j);
                        q[j] = v.x; q[j + 1] = v.y; q[j + 2] = v.z; q[j + 3] = v.w;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < CH; j++) q[j] = gp[j];
                }
                const float x1 = fmaf(l0.x, t - l0.y, l0.z), x2 = fmaf(l0.w, t - l1.x, l1.y);
                // the forward projector's IntegralBoxX weight of this pixel (dd_line)
                const float xl = min_f(x1, x2), xh = max_f(x1, x2);
                const float a = med_clamp(xl, 0.f, (float)n), b = med_clamp(xh, 0.f, (float)n);
                const float wt = (fminf(b, lo + 1.f) - fmaxf(a, lo)) * fast_div(l1.z, xh - xl);
                const float w = (b > a && wt > 0.f) ? wt : 0.f;
#pragma unroll
                for (int j = 0; j < CH; j++) acc[j] = fmaf(w, q[j], acc[j]);
            }
        }
    }
    store_bp_result<CH>(img, part, acc, grp, c0, C, (size_t)g.nx * g.ny, (size_t)iy * g.nx + ix);
}

struct DDAdjSetup {
    int C = 0, CH = 1, nchunk = 1, cs = 0, groups = 1, vpg = 1;
    bool rev = false;
    std::vector<int> cstart;
    std::vector<ChanRow> crow;
};

static DDAdjSetup dd_adj_setup(int B, const Grid& g, const Det& d, int nview, int type, bool fan) {
    const bool branchless = (type & DD_BRANCHLESS) != 0;
    DDAdjSetup s;
    s.C = B * g.nz;
    s.CH = pick_ch(s.C);
    s.nchunk = ceil_div(s.C, s.CH);
    s.cs = s.nchunk * s.CH;
    s.rev = !fan && (type & DD_FORCE_FBP) != 0 && !branchless;
    // channel -> (detector row, weight) lists: the transpose of the forward
    // projector's row table (same z mapping as dd2d_fp)
    std::vector<std::vector<ChanRow>> per(s.C);
    for (int b = 0; b < B; b++)
        for (int iv = 0; iv < d.nv; iv++) {
            const double zr = ((double)iv - (d.nv - 1) / 2.0 + d.off_v) * d.dv;
            double z = (zr - g.cz) / g.dz + g.nz / 2.0 - 0.5;
            z = std::min((double)g.nz, std::max(0.0, z));
            int iz = (int)z, iz1 = iz + 1;
            iz = clampi(iz, 0, g.nz);
            iz1 = clampi(iz1, 0, g.nz);
            const float w1 = (float)(z - iz);
            per[b * g.nz + iz].push_back(ChanRow{iv, 1.f - w1});
            per[b * g.nz + iz1].push_back(ChanRow{iv, w1});
        }
    s.cstart.assign(s.C + 1, 0);
    for (int c = 0; c < s.C; c++) {
        s.cstart[c] = (int)s.crow.size();
        s.crow.insert(s.crow.end(), per[c].begin(), per[c].end());
    }
    s.cstart[s.C] = (int)s.crow.size();
    s.groups = bp_view_groups((size_t)g.nx * g.ny, s.nchunk, nview);
    s.vpg = ceil_div(nview, s.groups);
    return s;
}

static void dd_adj_lines(FPLine* dlines, const DDView* dviews, const Grid& g, const Det& d, int nview, float dsd,
                         float dso, bool fan, hipStream_t stream) {
    if (fan) dd_fp_lines_kernel<true><<<dim3(ceil_div(d.nu, 256), nview), 256, 0, stream>>>(dlines, dviews, g, d, nview, dsd, dso);
    else dd_fp_lines_kernel<false><<<dim3(ceil_div(d.nu, 256), nview), 256, 0, stream>>>(dlines, dviews, g, d, nview, dsd, dso);
    CTP_LAUNCH_CHECK();
}

// The adjoint's launches after the line table: G is scratch of
// nview * nu * cs floats, part of groups * C * npix floats when groups > 1.
static void dd_adj_run(const DDAdjSetup& s, const DDView* dviews, const FPLine* dlines, const int* dcstart,
                       const ChanRow* dcrow, float* G, float* part, float* img, const float* prj, const Grid& g,
                       const Det& d, int nview, float dso, bool fan, hipStream_t stream) {
    const int C = s.C, CH = s.CH, nchunk = s.nchunk, cs = s.cs, groups = s.groups, vpg = s.vpg;
    dd_adj_rows_kernel<<<dim3(ceil_div(d.nu, 64), ceil_div(cs, 4), nview), 256, 0, stream>>>(
        G, prj, dviews, dcstart, dcrow, C, cs, g.nz, d, nview, s.rev);
    CTP_LAUNCH_CHECK();
    dim3 grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), nchunk * groups), block(256);
#define CTP_ADJ_CASE(N)                                                                                         \
    case N:                                                                                                     \
        if (fan) dd_fp_adjoint_kernel<N, true><<<grid, block, 0, stream>>>(img, part, G, dlines, dviews, C, cs, nchunk, \
                                                                           vpg, g, d, nview, dso);              \
        else dd_fp_adjoint_kernel<N, false><<<grid, block, 0, stream>>>(img, part, G, dlines, dviews, C, cs, nchunk, \
                                                                        vpg, g, d, nview, dso);                 \
        break;
    switch (CH) {
        CTP_ADJ_CASE(1) CTP_ADJ_CASE(2) CTP_ADJ_CASE(4) CTP_ADJ_CASE(8) CTP_ADJ_CASE(16) CTP_ADJ_CASE(32)
        default: fail("dd adjoint: bad channel chunk");
    }
#undef CTP_ADJ_CASE
    CTP_LAUNCH_CHECK();
    if (part) reduce_parts(img, part, groups, (size_t)C * g.nx * g.ny, stream);
}

void dd2d_fp_adjoint(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d,
                     int nview, float dsd, float dso, int type, bool fan, hipStream_t stream) {
    const std::vector<DDView> views = dd_views_for(deg_host, nview, g, type);
    const DDAdjSetup s = dd_adj_setup(B, g, d, nview, type, fan);
    const size_t npix = (size_t)g.nx * g.ny;
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<FPLine>((size_t)nview * d.nu);
    pl.add<float>((size_t)nview * d.nu * s.cs);
    pl.add<DDView>(nview);
    pl.add<int>(s.C + 1);
    pl.add<ChanRow>(std::max<size_t>(s.crow.size(), 1));
    if (s.groups > 1) pl.add<float>((size_t)s.groups * s.C * npix);
    ws.reserve(pl.bytes);
    FPLine* dlines = ws.take_n<FPLine>((size_t)nview * d.nu);
    float* G = ws.take_n<float>((size_t)nview * d.nu * s.cs);
    DDView* dviews = ws.take_n<DDView>(nview);
    int* dcstart = ws.take_n<int>(s.C + 1);
    ChanRow* dcrow = ws.take_n<ChanRow>(std::max<size_t>(s.crow.size(), 1));
    float* part = s.groups > 1 ? ws.take_n<float>((size_t)s.groups * s.C * npix) : nullptr;
    CTP_CHECK(hipMemcpyAsync(dviews, views.data(), sizeof(DDView) * nview, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dcstart, s.cstart.data(), sizeof(int) * (s.C + 1), hipMemcpyHostToDevice, stream));
    if (!s.crow.empty())
        CTP_CHECK(hipMemcpyAsync(dcrow, s.crow.data(), sizeof(ChanRow) * s.crow.size(), hipMemcpyHostToDevice, stream));
    TimedLaunch t(fan ? "dd_fan_fp_adjoint" : "dd_par_fp_adjoint", stream);
    dd_adj_lines(dlines, dviews, g, d, nview, dsd, dso, fan, stream);
    dd_adj_run(s, dviews, dlines, dcstart, dcrow, G, part, img, prj, g, d, nview, dso, fan, stream);
    CTP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// persistent plans (no counterpart in the reference)
// ---------------------------------------------------------------------------
// One geometry and batch count: every geometry-only table of the forward
// projector, the backprojector and the exact adjoint is built once and kept on
// the device, so that the calls launch kernels only (no host work, no
// device-to-host angle copy, no synchronisation).  Scratch comes from the
// calling stream's workspace.
struct DDPlan {
    int device = 0;
    int B = 0, nview = 0, type = 0;
    Grid g{};
    Det d{};
    float dsd = 0.f, dso = 0.f;
    bool fan = true;
    DDFPSetup fp;
    DDBPSetup bp;
    DDAdjSetup adj;
    void* mem = nullptr;
    DDView* views = nullptr;
    FPRow* rows = nullptr;
    int* start = nullptr;
    FPGroup* groups = nullptr;
    SliceRow* srow = nullptr;
    int* cstart = nullptr;
    ChanRow* crow = nullptr;
    FPLine* lines = nullptr;
    // scratch of each direction, owned by the plan (no shared workspace, so
    // that calls can be captured in a HIP graph): channel-last image copies
    // of the forward, channel-last sinogram and partial images of the
    // backprojection and of the adjoint
    void* scratch = nullptr;
    float *imgY = nullptr, *imgX = nullptr;
    float *bp_sino = nullptr, *bp_part = nullptr;
    float *adj_G = nullptr, *adj_part = nullptr;
    ~DDPlan() {
        if (mem) (void)hipFree(mem);
        if (scratch) (void)hipFree(scratch);
    }
};

DDPlan* dd_plan_create(const float* deg_host, int B, const Grid& g, const Det& d, int nview, float dsd, float dso,
                       int type, bool fan, hipStream_t stream) {
    std::unique_ptr<DDPlan> p(new DDPlan());
    CTP_CHECK(hipGetDevice(&p->device));
    p->B = B; p->nview = nview; p->type = type; p->g = g; p->d = d; p->dsd = dsd; p->dso = dso; p->fan = fan;
    const std::vector<DDView> views = dd_views_for(deg_host, nview, g, type);
    p->fp = dd_fp_setup(views, B, g, d, nview, dsd, dso, type, fan);
    p->bp = dd_bp_setup(B, g, d, nview, type);
    p->adj = dd_adj_setup(B, g, d, nview, type, fan);
    Plan pl;
    pl.add<DDView>(nview);
    pl.add<FPRow>(std::max<size_t>(p->fp.rows.size(), 1));
    pl.add<int>(p->fp.nchunk + 1);
    pl.add<FPGroup>(std::max<size_t>(p->fp.groups.size(), 1));
    pl.add<SliceRow>(g.nz);
    pl.add<int>(p->adj.C + 1);
    pl.add<ChanRow>(std::max<size_t>(p->adj.crow.size(), 1));
    pl.add<FPLine>((size_t)nview * d.nu);
    CTP_CHECK(hipMalloc(&p->mem, pl.bytes));
    char* m = static_cast<char*>(p->mem);
    auto carve = [&m](size_t bytes) { void* r = m; m += (bytes + 255) / 256 * 256; return r; };
    p->views = static_cast<DDView*>(carve(sizeof(DDView) * nview));
    p->rows = static_cast<FPRow*>(carve(sizeof(FPRow) * std::max<size_t>(p->fp.rows.size(), 1)));
    p->start = static_cast<int*>(carve(sizeof(int) * (p->fp.nchunk + 1)));
    p->groups = static_cast<FPGroup*>(carve(sizeof(FPGroup) * std::max<size_t>(p->fp.groups.size(), 1)));
    p->srow = static_cast<SliceRow*>(carve(sizeof(SliceRow) * g.nz));
    p->cstart = static_cast<int*>(carve(sizeof(int) * (p->adj.C + 1)));
    p->crow = static_cast<ChanRow*>(carve(sizeof(ChanRow) * std::max<size_t>(p->adj.crow.size(), 1)));
    p->lines = static_cast<FPLine*>(carve(sizeof(FPLine) * (size_t)nview * d.nu));
    CTP_CHECK(hipMemcpyAsync(p->views, views.data(), sizeof(DDView) * nview, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(p->rows, p->fp.rows.data(), sizeof(FPRow) * p->fp.rows.size(), hipMemcpyHostToDevice,
                             stream));
    CTP_CHECK(hipMemcpyAsync(p->start, p->fp.start.data(), sizeof(int) * (p->fp.nchunk + 1), hipMemcpyHostToDevice,
                             stream));
    CTP_CHECK(hipMemcpyAsync(p->groups, p->fp.groups.data(), sizeof(FPGroup) * p->fp.groups.size(),
                             hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(p->srow, p->bp.srow.data(), sizeof(SliceRow) * g.nz, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(p->cstart, p->adj.cstart.data(), sizeof(int) * (p->adj.C + 1), hipMemcpyHostToDevice,
                             stream));
    if (!p->adj.crow.empty())
        CTP_CHECK(hipMemcpyAsync(p->crow, p->adj.crow.data(), sizeof(ChanRow) * p->adj.crow.size(),
                                 hipMemcpyHostToDevice, stream));
    dd_adj_lines(p->lines, p->views, g, d, nview, dsd, dso, fan, stream);
    const size_t npix = (size_t)g.nx * g.ny;
    Plan ps;
    ps.add<float>((size_t)g.ny * p->fp.rsY);
    ps.add<float>((size_t)g.nx * p->fp.rsX);
    ps.add<float>((size_t)nview * d.nu * p->bp.cs);
    ps.add<float>(p->bp.groups > 1 ? (size_t)p->bp.groups * p->bp.C * npix : 0);
    ps.add<float>((size_t)nview * d.nu * p->adj.cs);
    ps.add<float>(p->adj.groups > 1 ? (size_t)p->adj.groups * p->adj.C * npix : 0);
    CTP_CHECK(hipMalloc(&p->scratch, ps.bytes));
    char* q = static_cast<char*>(p->scratch);
    auto carve_f = [&q](size_t n) { float* r = reinterpret_cast<float*>(q); q += (n * sizeof(float) + 255) / 256 * 256; return r; };
    p->imgY = carve_f((size_t)g.ny * p->fp.rsY);
    p->imgX = carve_f((size_t)g.nx * p->fp.rsX);
    p->bp_sino = carve_f((size_t)nview * d.nu * p->bp.cs);
    p->bp_part = p->bp.groups > 1 ? carve_f((size_t)p->bp.groups * p->bp.C * npix) : nullptr;
    p->adj_G = carve_f((size_t)nview * d.nu * p->adj.cs);
    p->adj_part = p->adj.groups > 1 ? carve_f((size_t)p->adj.groups * p->adj.C * npix) : nullptr;
    CTP_CHECK(hipStreamSynchronize(stream));
    return p.release();
}

static void dd_plan_check(const DDPlan* p) {
    int dev = -1;
    CTP_CHECK(hipGetDevice(&dev));
    if (dev != p->device) fail("dd plan: called on another device than the one it was created on");
}

void dd_plan_forward(DDPlan* p, float* prj, const float* img, hipStream_t stream) {
    dd_plan_check(p);
    dd_fp_run(p->fp, DDFPDev{p->views, p->rows, p->start, p->groups}, prj, img, p->imgY, p->imgX, p->g, p->d,
              p->nview, p->dsd, p->dso, p->fan, stream);
}

void dd_plan_backward(DDPlan* p, float* img, const float* prj, hipStream_t stream) {
    dd_plan_check(p);
    dd_bp_run(p->bp, p->views, p->srow, p->bp_sino, p->bp_part, img, prj, p->g, p->d, p->nview, p->dso, p->fan,
              stream);
}

void dd_plan_adjoint(DDPlan* p, float* img, const float* prj, hipStream_t stream) {
    dd_plan_check(p);
    TimedLaunch t(p->fan ? "dd_fan_fp_adjoint" : "dd_par_fp_adjoint", stream);
    dd_adj_run(p->adj, p->views, p->lines, p->cstart, p->crow, p->adj_G, p->adj_part, img, prj, p->g, p->d,
               p->nview, p->dso, p->fan, stream);
}

void dd_plan_destroy(DDPlan* p) {
    if (!p) return;
    std::unique_ptr<DDPlan> owned(p);
    int cur = 0;
    CTP_CHECK(hipGetDevice(&cur));
    CTP_CHECK(hipSetDevice(p->device));
    CTP_CHECK(hipDeviceSynchronize());   // queued calls may still read the tables
    owned.reset();
    CTP_CHECK(hipSetDevice(cur));
}


// ---------------------------------------------------------------------------
// batched plans of uneven jobs (no counterpart in the reference)
// ---------------------------------------------------------------------------
// Many fan / parallel projections of different geometry, size, view count and
// batch count.  Each job keeps the launch shapes and tables a single call of
// its own takes (so its results are those of the single call), and all jobs
// whose kernel is the same template instance run in one launch: a block finds
// its job by a binary search over the jobs' first block indices and runs the
// single-call kernel's block body with that job's parameters.  Forward: one
// launch making the channel-last copies of every job, then one per forward
// kernel class.  Backward: one launch clearing every image, one making the
// weighted channel-last projections of every job, one per backprojector
// class, one summing the partial images.  Blocks take their index from a work
// counter (persistent blocks) or from the hardware dispatcher.  The calls only
// launch kernels (counters are cleared by a kernel), so they can be captured
// in a HIP graph.

struct DDJobDev {
    float* prj;
    float* img;
    float* imgY;
    float* imgX;
    float* sino;
    float* part;
    const DDView* views;
    const FPRow* rows;
    const int* start;
    const SliceRow* srow;
    long rsY, rsX;
    Grid g;
    Det d;
    int nview, fan;
    float dsd, dso;
    int fC, fcs, nrow, lseg, force, needY, needX;   // forward
    int bC, bcs, nchunk, groups, vpg, fbp;          // backprojection
    int tr_gx, tr_gy;      // channel-last copies: tr_gx x tr_gy x ny blocks per copy
    int fp_gx, fp_gy;      // forward projection (third grid dimension: channel chunks)
    int pr_gx, pr_gy;      // weighted projections (third: views)
    int bp_gx, bp_gy;      // backprojection (third: chunks x view groups)
};

// Member k of a launch whose blocks begin[k] .. begin[k + 1] - 1 are block b's
// (begin ascending from 0; members without blocks repeat the next begin).
__device__ inline int dd_member(const int* __restrict__ begin, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (const_load(begin, mid) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define DD_BATCH_ARGS                                                                                         \
    const DDJobDev *__restrict__ jobs, const int *__restrict__ members, const int *__restrict__ begin, int n, \
        int total, int *__restrict__ ctr, int persist
#define DD_BATCH_LOOP                                                                                          \
    __shared__ int slot;                                                                                       \
    for (int b = first_block(ctr, 1, &slot, persist); b < total; b = following_block(ctr, 1, &slot, persist))
#define DD_BATCH_JOB                                                                                           \
    const int k = dd_member(begin, n, b);                                                                      \
    const DDJobDev J = const_load(jobs, (long)const_load(members, k));                                         \
    const int l = b - const_load(begin, k)

__global__ __launch_bounds__(256) void dd_batch_copy_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        // the y-major copy (if needed), then the x-major copy (if needed)
        const int per = J.tr_gx * J.tr_gy * J.g.ny;
        const bool xmajor = !J.needY || l >= per;
        const int m = (J.needY && xmajor) ? l - per : l;
        const int bx = m % J.tr_gx, t = m / J.tr_gx;
        to_channel_last_block(xmajor ? J.imgX : J.imgY, J.img, J.fC, J.fcs, J.g.nx, J.g.ny, xmajor,
                              (size_t)(xmajor ? J.rsX : J.rsY), bx, t % J.tr_gy, t / J.tr_gy);
    }
}

template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void dd_batch_fp_split_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        dd_fp_split_block<CH, LPC, FAN>(J.prj, J.imgY, J.imgX, J.views, J.rows, J.nrow, J.g, J.d, J.nview, J.dsd,
                                        J.dso, J.force != 0, l % J.fp_gx, l / J.fp_gx);
    }
}

template <int CH, bool FAN, bool XC, int VPB>
__global__ __launch_bounds__(VPB > 1 ? 64 * VPB : 128) void dd_batch_fp_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const int t = l / J.fp_gx;
        dd_fp_block<CH, FAN, XC, VPB>(J.prj, J.imgY, J.imgX, J.views, J.rows, J.start, J.fcs, J.g, J.d, J.nview,
                                      J.dsd, J.dso, J.force != 0, J.lseg, l % J.fp_gx, t % J.fp_gy, t / J.fp_gy);
    }
}

// 1024 floats of one image per block
__global__ __launch_bounds__(256) void dd_batch_zero_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const size_t nimg = (size_t)J.bC * J.g.nx * J.g.ny;
        const size_t i0 = (size_t)l * 1024 + threadIdx.x;
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (i0 + 256 * e < nimg) J.img[i0 + 256 * e] = 0.f;
    }
}

__global__ __launch_bounds__(256) void dd_batch_prep_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const int t = l / J.pr_gx;
        dd_bp_prep_block(J.sino, J.prj, J.views, J.srow, J.bC, J.bcs, J.g.nz, J.d, J.nview, J.g.dx, J.g.dy,
                         J.fan != 0, J.fbp == 0, l % J.pr_gx, t % J.pr_gy, t / J.pr_gy);
    }
}

template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void dd_batch_bp_split_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const int t = l / J.bp_gx;
        dd_bp_split_block<CH, LPC, FAN>(J.img, J.part, J.sino, J.views, J.bC, J.bcs, J.nchunk, J.vpg, J.g, J.d,
                                        J.nview, J.dso, J.fbp != 0, l % J.bp_gx, t % J.bp_gy, t / J.bp_gy);
    }
}

template <int CH, bool FAN>
__global__ __launch_bounds__(256) void dd_batch_bp_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const int t = l / J.bp_gx;
        dd_bp_block<CH, FAN>(J.img, J.part, J.sino, J.views, J.bC, J.bcs, J.nchunk, J.vpg, J.g, J.d, J.nview, J.dso,
                             J.fbp != 0, l % J.bp_gx, t % J.bp_gy, t / J.bp_gy);
    }
}

// the view groups' partial images summed in group order (as reduce_parts)
__global__ __launch_bounds__(256) void dd_batch_reduce_kernel(DD_BATCH_ARGS) {
    DD_BATCH_LOOP {
        DD_BATCH_JOB;
        const size_t nimg = (size_t)J.bC * J.g.nx * J.g.ny;
        const size_t i = (size_t)l * 256 + threadIdx.x;
        if (i < nimg) {
            float s = 0.f;
            for (int q = 0; q < J.groups; q++) s += J.part[(size_t)q * nimg + i];
            J.img[i] += s;
        }
    }
}

#undef DD_BATCH_JOB
#undef DD_BATCH_LOOP
#undef DD_BATCH_ARGS

__global__ void dd_batch_clear_kernel(int* __restrict__ ctr, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) ctr[i] = 0;
}

constexpr int kDDBatchMaxFanout = 4;

namespace {

enum DDLaunchKind { DDL_COPY, DDL_FP_SPLIT, DDL_FP, DDL_ZERO, DDL_PREP, DDL_BP_SPLIT, DDL_BP, DDL_REDUCE };

struct DDLaunch {
    int kind = DDL_COPY;
    int ch = 1, lpc = 1, vpb = 1;
    bool fan = true, xc = false;
    std::vector<int> members, begin;   // begin: first block of each member, then the total
    int* dmembers = nullptr;
    int* dbegin = nullptr;
    int total = 0, block = 256, grid = 0;
    int ctr = 0;                       // index of the launch's work counter
    bool same_kernel(const DDLaunch& o) const {
        return kind == o.kind && ch == o.ch && lpc == o.lpc && vpb == o.vpb && fan == o.fan && xc == o.xc;
    }
};

}  // namespace

struct DDBatchPlan::Impl {
    int device = 0;
    std::vector<DDBatchJobSpec> spec;
    std::vector<std::vector<DDView>> views;
    std::vector<DDFPSetup> fps;
    std::vector<DDBPSetup> bps;
    std::vector<DDJobDev> jobs;
    DDJobDev* djobs = nullptr;
    std::vector<DDLaunch> fwd, bwd;
    int* ctr = nullptr;                // work counters: forward launches, then backward launches
    int nctr = 0;
    int fp_lpc = 0, bp_lpc = 0;        // 0: the single calls' choice (tuning keys)
    int pf = 1, pb = 1;                // persistent blocks (1) or hardware dispatch (0), per direction
    std::vector<void*> mem;            // job buffers and tables
    std::vector<void*> lmem;           // launch tables of the current layout
    int fanout_f = 1, fanout_b = 1;    // streams the kernel classes of each direction run on
    std::vector<hipStream_t> side;     // side streams (created with the plan: none may be created under capture)
    std::vector<hipEvent_t> events;    // fork, then one join per side stream

    template <typename T>
    T* alloc(std::vector<void*>& owner, size_t n) {
        void* p = nullptr;
        CTP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        owner.push_back(p);
        return static_cast<T*>(p);
    }
    ~Impl() {
        for (auto st : side) (void)hipStreamSynchronize(st);
        for (void* p : lmem) (void)hipFree(p);
        for (void* p : mem) (void)hipFree(p);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto st : side) (void)hipStreamDestroy(st);
    }
    void setup_jobs();
    void launch(const DDLaunch& L, int persist, hipStream_t s);
    void layout();
    void run(const std::vector<DDLaunch>& ls, int ctr_base, int persist, int fanout, hipStream_t s,
             const char* what);
    float time_direction(bool forward_dir, hipStream_t s);
};

// per job: tables and scratch as a single call makes them (on the device for
// the plan's lifetime); the channel-last copies are sized for the widest
// pixel stride any lanes-per-column choice gives
void DDBatchPlan::Impl::setup_jobs() {
    const int nj = (int)spec.size();
    views.resize(nj);
    jobs.assign(nj, DDJobDev{});
    for (int j = 0; j < nj; j++) {
        const DDBatchJobSpec& S = spec[j];
        views[j] = dd_views_for(S.deg, S.nview, S.g, S.type);
        const DDFPSetup widest = dd_fp_setup(views[j], S.B, S.g, S.d, S.nview, S.dsd, S.dso, S.type, S.fan, false, 1);
        const DDFPSetup fs = dd_fp_setup(views[j], S.B, S.g, S.d, S.nview, S.dsd, S.dso, S.type, S.fan, false, 0);
        const DDBPSetup bs = dd_bp_setup(S.B, S.g, S.d, S.nview, S.type);
        DDJobDev& J = jobs[j];
        J.prj = S.prj;
        J.img = S.img;
        J.g = S.g;
        J.d = S.d;
        J.nview = S.nview;
        J.fan = S.fan ? 1 : 0;
        J.dsd = S.dsd;
        J.dso = S.dso;
        const size_t ny = S.g.ny, nx = S.g.nx, npix = nx * ny;
        const long rs_max = std::max(widest.rsY, fs.rsY), rx_max = std::max(widest.rsX, fs.rsX);
        J.imgY = fs.needY ? alloc<float>(mem, ny * rs_max) : nullptr;
        J.imgX = fs.needX ? alloc<float>(mem, nx * rx_max) : nullptr;
        J.sino = alloc<float>(mem, (size_t)S.nview * S.d.nu * bs.cs);
        J.part = bs.groups > 1 ? alloc<float>(mem, (size_t)bs.groups * bs.C * npix) : nullptr;
        DDView* dv = alloc<DDView>(mem, S.nview);
        FPRow* dr = alloc<FPRow>(mem, fs.rows.size());
        int* dst = alloc<int>(mem, fs.nchunk + 1);
        SliceRow* dsr = alloc<SliceRow>(mem, S.g.nz);
        CTP_CHECK(hipMemcpy(dv, views[j].data(), sizeof(DDView) * S.nview, hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(dr, fs.rows.data(), sizeof(FPRow) * fs.rows.size(), hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(dst, fs.start.data(), sizeof(int) * (fs.nchunk + 1), hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(dsr, bs.srow.data(), sizeof(SliceRow) * S.g.nz, hipMemcpyHostToDevice));
        J.views = dv;
        J.rows = dr;
        J.start = dst;
        J.srow = dsr;
        bps.push_back(bs);
    }
}

// Kernel of a launch (nullptr: no such instance).
static const void* dd_launch_kernel(const DDLaunch& L) {
    switch (L.kind) {
    case DDL_COPY: return reinterpret_cast<const void*>(dd_batch_copy_kernel);
    case DDL_ZERO: return reinterpret_cast<const void*>(dd_batch_zero_kernel);
    case DDL_PREP: return reinterpret_cast<const void*>(dd_batch_prep_kernel);
    case DDL_REDUCE: return reinterpret_cast<const void*>(dd_batch_reduce_kernel);
    case DDL_FP_SPLIT:
#define CTP_K(N, P)                                                                                            \
    if (L.ch == N && L.lpc == P)                                                                               \
        return L.fan ? reinterpret_cast<const void*>(dd_batch_fp_split_kernel<N, P, true>)                     \
                     : reinterpret_cast<const void*>(dd_batch_fp_split_kernel<N, P, false>);
        CTP_K(8, 2) CTP_K(16, 2) CTP_K(16, 4) CTP_K(32, 2) CTP_K(32, 4) CTP_K(32, 8)
#undef CTP_K
        return nullptr;
    case DDL_FP:
#define CTP_K(N)                                                                                               \
    if (L.ch == N) {                                                                                           \
        if (L.vpb == 4) return L.fan ? (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, true, 4>) \
                                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, false, 4>)) \
                                     : (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, true, 4>) \
                                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, false, 4>)); \
        if (L.vpb == 2) return L.fan ? (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, true, 2>) \
                                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, false, 2>)) \
                                     : (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, true, 2>) \
                                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, false, 2>)); \
        return L.fan ? (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, true, 1>)             \
                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, true, false, 1>))           \
                     : (L.xc ? reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, true, 1>)            \
                             : reinterpret_cast<const void*>(dd_batch_fp_kernel<N, false, false, 1>));         \
    }
        CTP_K(1) CTP_K(2) CTP_K(4) CTP_K(8) CTP_K(16) CTP_K(32)
#undef CTP_K
        return nullptr;
    case DDL_BP_SPLIT:
#define CTP_K(N, P)                                                                                            \
    if (L.ch == N && L.lpc == P)                                                                               \
        return L.fan ? reinterpret_cast<const void*>(dd_batch_bp_split_kernel<N, P, true>)                     \
                     : reinterpret_cast<const void*>(dd_batch_bp_split_kernel<N, P, false>);
        CTP_K(8, 2) CTP_K(16, 2) CTP_K(16, 4) CTP_K(32, 2) CTP_K(32, 4)
#undef CTP_K
        return nullptr;
    case DDL_BP:
#define CTP_K(N)                                                                                               \
    if (L.ch == N)                                                                                             \
        return L.fan ? reinterpret_cast<const void*>(dd_batch_bp_kernel<N, true>)                              \
                     : reinterpret_cast<const void*>(dd_batch_bp_kernel<N, false>);
        CTP_K(1) CTP_K(2) CTP_K(4) CTP_K(8) CTP_K(16) CTP_K(32)
#undef CTP_K
        return nullptr;
    }
    return nullptr;
}

// Launch groups and the job records of the current choice of lanes per column
// (forward), lanes per pixel (backward) and block distribution.
void DDBatchPlan::Impl::layout() {
    for (void* p : lmem) (void)hipFree(p);
    lmem.clear();
    fwd.clear();
    bwd.clear();
    fps.clear();
    const int nj = (int)spec.size();
    const int tv = tuning_get("dd_fp_views_per_block", 4);
    const long limit = 0x7fffffffL - (1L << 24);   // block indices (and counters running past them) in 31 bits
    auto add = [&](std::vector<DDLaunch>& ls, DDLaunch L, int j, long blocks) {
        if (blocks <= 0) return;
        DDLaunch* at = nullptr;
        for (auto& o : ls)
            if (o.same_kernel(L)) at = &o;
        if (!at) {
            ls.push_back(L);
            at = &ls.back();
        }
        if ((long)at->total + blocks > limit) fail("dd batch: more than 2^31 blocks in one launch");
        at->members.push_back(j);
        at->begin.push_back(at->total);
        at->total += (int)blocks;
    };
    for (int j = 0; j < nj; j++) {
        const DDBatchJobSpec& S = spec[j];
        DDJobDev& J = jobs[j];
        const DDFPSetup fs = dd_fp_setup(views[j], S.B, S.g, S.d, S.nview, S.dsd, S.dso, S.type, S.fan, false, fp_lpc);
        fps.push_back(fs);
        const DDBPSetup& bs = bps[j];
        J.fC = fs.C; J.fcs = fs.cs; J.nrow = (int)fs.rows.size(); J.lseg = fs.lseg; J.force = fs.force ? 1 : 0;
        J.needY = fs.needY ? 1 : 0; J.needX = fs.needX ? 1 : 0;
        J.rsY = fs.rsY; J.rsX = fs.rsX;
        J.bC = bs.C; J.bcs = bs.cs; J.nchunk = bs.nchunk; J.groups = bs.groups; J.vpg = bs.vpg;
        J.fbp = bs.fbp ? 1 : 0;
        // forward
        J.tr_gx = ceil_div(S.g.nx, 32);
        J.tr_gy = ceil_div(fs.cs, 32);
        DDLaunch c;
        c.kind = DDL_COPY;
        add(fwd, c, j, (long)(J.needY + J.needX) * J.tr_gx * J.tr_gy * S.g.ny);
        DDLaunch f;
        f.fan = S.fan;
        f.ch = fs.CH;
        if (fs.lpc > 1) {
            f.kind = DDL_FP_SPLIT;
            f.lpc = fs.lpc;
            J.fp_gx = ceil_div(S.d.nu, 64 / fs.lpc);
            J.fp_gy = ceil_div(S.nview, 4);
            add(fwd, f, j, (long)J.fp_gx * J.fp_gy);
        } else {
            f.kind = DDL_FP;
            f.xc = fs.xc;
            f.vpb = (fs.lseg == 0 && (tv == 2 || tv == 4)) ? tv : 1;
            f.block = f.vpb > 1 ? 64 * f.vpb : 128;
            J.fp_gx = f.vpb > 1 ? ceil_div(S.d.nu, 64) : ceil_div(S.d.nu, 128 >> fs.lseg);
            J.fp_gy = f.vpb > 1 ? ceil_div(S.nview, f.vpb) : S.nview;
            add(fwd, f, j, (long)J.fp_gx * J.fp_gy * fs.nchunk);
        }
        // backward
        const long nimg = (long)bs.C * S.g.nx * S.g.ny;
        DDLaunch z;
        z.kind = DDL_ZERO;
        add(bwd, z, j, (nimg + 1023) / 1024);
        J.pr_gx = ceil_div(S.d.nu, 32);
        J.pr_gy = ceil_div(bs.cs, 32);
        DDLaunch pr;
        pr.kind = DDL_PREP;
        add(bwd, pr, j, (long)J.pr_gx * J.pr_gy * S.nview);
        int lpc = bp_lpc > 0 ? (bp_lpc == 4 && bs.CH % 16 != 0 ? 2 : bp_lpc) : dd_bp_lanes(bs.CH);
        DDLaunch bk;
        bk.fan = S.fan;
        bk.ch = bs.CH;
        if ((lpc == 2 || lpc == 4) && bs.CH >= 8 && bs.CH % (4 * lpc) == 0) {
            bk.kind = DDL_BP_SPLIT;
            bk.lpc = lpc;
            J.bp_gx = ceil_div(S.g.nx, lpc == 2 ? 16 : 8);
            J.bp_gy = ceil_div(S.g.ny, 8);
        } else {
            bk.kind = DDL_BP;
            J.bp_gx = ceil_div(S.g.nx, 16);
            J.bp_gy = ceil_div(S.g.ny, 16);
        }
        add(bwd, bk, j, (long)J.bp_gx * J.bp_gy * bs.nchunk * bs.groups);
        if (bs.groups > 1) {
            DDLaunch r;
            r.kind = DDL_REDUCE;
            add(bwd, r, j, (nimg + 255) / 256);
        }
    }
    // stream order: copies before projections; cleared images before the
    // weighted projections, the backprojections and the partial-image sums
    auto by_kind = [](const DDLaunch& x, const DDLaunch& y) { return x.kind < y.kind; };
    std::stable_sort(fwd.begin(), fwd.end(), by_kind);
    std::stable_sort(bwd.begin(), bwd.end(), by_kind);
    int ci = 0;
    for (auto* ls : {&fwd, &bwd})
        for (auto& L : *ls) {
            const int persist = ls == &fwd ? pf : pb;
            L.begin.push_back(L.total);
            L.dmembers = alloc<int>(lmem, L.members.size());
            L.dbegin = alloc<int>(lmem, L.begin.size());
            CTP_CHECK(hipMemcpy(L.dmembers, L.members.data(), sizeof(int) * L.members.size(), hipMemcpyHostToDevice));
            CTP_CHECK(hipMemcpy(L.dbegin, L.begin.data(), sizeof(int) * L.begin.size(), hipMemcpyHostToDevice));
            const void* k = dd_launch_kernel(L);
            if (!k) fail("dd batch: no kernel for a job's launch shape");
            if (persist) {
                int dev = 0, nb = 0, ncu = 0;
                CTP_CHECK(hipGetDevice(&dev));
                CTP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
                CTP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, L.block, 0));
                L.grid = std::max(1, std::min(L.total, std::max(nb, 1) * ncu));
            } else {
                L.grid = L.total;
            }
            L.ctr = ci++;
        }
    nctr = ci;
    ctr = alloc<int>(lmem, std::max(nctr, 1));
    djobs = alloc<DDJobDev>(lmem, jobs.size());
    CTP_CHECK(hipMemcpy(djobs, jobs.data(), sizeof(DDJobDev) * jobs.size(), hipMemcpyHostToDevice));
}

void DDBatchPlan::Impl::launch(const DDLaunch& L, int persist, hipStream_t s) {
    const void* k = dd_launch_kernel(L);
    int* c = ctr + L.ctr;
    int n = (int)L.members.size();
    int total = L.total;
    int pers = persist;
    const DDJobDev* jb = djobs;
    const int* mb = L.dmembers;
    const int* bg = L.dbegin;
    void* args[] = {(void*)&jb, (void*)&mb, (void*)&bg, (void*)&n, (void*)&total, (void*)&c, (void*)&pers};
    CTP_CHECK(hipLaunchKernel(k, dim3(L.grid), dim3(L.block), args, 0, s));
}

// The projection / backprojection classes of one direction touch disjoint jobs,
// so with fanout > 1 they run on side streams forked from 's' after the copies
// (weighted projections) and joined before the partial-image sums; the fork
// and join are events, which a graph capture turns into graph edges.
void DDBatchPlan::Impl::run(const std::vector<DDLaunch>& ls, int ctr_base, int persist, int fanout,
                            hipStream_t s, const char* what) {
    if (ls.empty()) return;
    if (persist) {
        dd_batch_clear_kernel<<<1, 64, 0, s>>>(ctr + ctr_base, (int)ls.size());
        CTP_LAUNCH_CHECK();
    }
    TimedLaunch t(what, s);
    std::vector<const DDLaunch*> mid;
    for (const DDLaunch& L : ls)
        if (L.kind == DDL_FP_SPLIT || L.kind == DDL_FP || L.kind == DDL_BP_SPLIT || L.kind == DDL_BP) mid.push_back(&L);
    const int nside = std::min<int>(std::min<int>(fanout, (int)side.size()), (int)mid.size());
    bool forked = false;
    for (const DDLaunch& L : ls) {
        const bool is_mid = std::find(mid.begin(), mid.end(), &L) != mid.end();
        if (!is_mid || nside <= 1) {
            if (forked) {   // join before the first launch after the classes
                for (int i = 0; i < nside; i++) {
                    CTP_CHECK(hipEventRecord(events[1 + i], side[i]));
                    CTP_CHECK(hipStreamWaitEvent(s, events[1 + i], 0));
                }
                forked = false;
            }
            launch(L, persist, s);
            continue;
        }
        if (!forked) {
            CTP_CHECK(hipEventRecord(events[0], s));
            for (int i = 0; i < nside; i++) CTP_CHECK(hipStreamWaitEvent(side[i], events[0], 0));
            forked = true;
        }
        const int i = (int)(std::find(mid.begin(), mid.end(), &L) - mid.begin()) % nside;
        launch(L, persist, side[i]);
    }
    if (forked)
        for (int i = 0; i < nside; i++) {
            CTP_CHECK(hipEventRecord(events[1 + i], side[i]));
            CTP_CHECK(hipStreamWaitEvent(s, events[1 + i], 0));
        }
}

float DDBatchPlan::Impl::time_direction(bool forward_dir, hipStream_t s) {
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
    float best = 1e30f;
    for (int rep = 0; rep < 4; rep++) {   // the first run warms up
        CTP_CHECK(hipEventRecord(a, s));
        if (forward_dir) run(fwd, 0, pf, fanout_f, s, "dd_batch_fp");
        else run(bwd, (int)fwd.size(), pb, fanout_b, s, "dd_batch_bp");
        CTP_CHECK(hipEventRecord(b, s));
        CTP_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        CTP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (rep > 0) best = std::min(best, ms);
    }
    return best;
}

DDBatchPlan::DDBatchPlan(const std::vector<DDBatchJobSpec>& jobs, bool autotune) : impl_(new Impl()) {
    Impl& I = *impl_;
    CTP_CHECK(hipGetDevice(&I.device));
    I.spec = jobs;
    // untuned defaults: hardware-dispatched blocks and the kernel classes on
    // four streams, what the autotuner chose for sixteen uneven jobs (1.9x
    // faster there than persistent blocks on one stream)
    I.pf = I.pb = tuning_get("dd_batch_persistent", 2) == 1 ? 1 : 0;
    I.fanout_f = I.fanout_b = std::max(1, std::min(tuning_get("dd_batch_streams", 4), kDDBatchMaxFanout));
    I.side.resize(kDDBatchMaxFanout);
    for (auto& st : I.side) CTP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    I.events.resize(kDDBatchMaxFanout + 1);
    for (auto& e : I.events) CTP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    I.setup_jobs();
    I.layout();
    if (autotune) tune();
}

DDBatchPlan::~DDBatchPlan() = default;

// Forward: lanes per detector column (2, 4, 8) of the split-lane kernel;
// backward: lanes per pixel (2, 4); each direction: persistent or
// hardware-dispatched blocks.  Each candidate is timed on the bound buffers
// (best of three runs after one warm-up); the results do not depend on the
// choice.
void DDBatchPlan::tune() {
    Impl& I = *impl_;
    TimingSuspended quiet;
    struct OwnStream {   // the timing stream, released on every exit
        hipStream_t s = nullptr;
        ~OwnStream() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        }
    } own;
    CTP_CHECK(hipStreamCreate(&own.s));
    hipStream_t s = own.s;
    float best_f = 1e30f, best_b = 1e30f;
    int bf = 4, bb = 4, bpf = 1, bpb = 1, bof = 1, bob = 1;
    for (int pers = 1; pers >= 0; pers--)
        for (int lpc : {2, 4, 8}) {
            I.pf = I.pb = pers;
            I.fp_lpc = lpc;
            I.bp_lpc = std::min(lpc, 4);
            I.layout();
            for (int fo : {1, kDDBatchMaxFanout}) {
                I.fanout_f = I.fanout_b = fo;
                const float tf = I.time_direction(true, s);
                if (tf < best_f) { best_f = tf; bf = lpc; bpf = pers; bof = fo; }
                if (lpc <= 4) {
                    const float tb = I.time_direction(false, s);
                    if (tb < best_b) { best_b = tb; bb = lpc; bpb = pers; bob = fo; }
                }
            }
        }
    I.fp_lpc = bf;
    I.bp_lpc = bb;
    I.pf = bpf;
    I.pb = bpb;
    I.fanout_f = bof;
    I.fanout_b = bob;
    I.layout();
    CTP_CHECK(hipStreamSynchronize(s));
}

void DDBatchPlan::forward(hipStream_t s) {
    int dev = -1;
    CTP_CHECK(hipGetDevice(&dev));
    if (dev != impl_->device) fail("dd batch: called on another device than the one it was created on");
    impl_->run(impl_->fwd, 0, impl_->pf, impl_->fanout_f, s, "dd_batch_fp");
}

void DDBatchPlan::backward(hipStream_t s) {
    int dev = -1;
    CTP_CHECK(hipGetDevice(&dev));
    if (dev != impl_->device) fail("dd batch: called on another device than the one it was created on");
    impl_->run(impl_->bwd, (int)impl_->fwd.size(), impl_->pb, impl_->fanout_b, s, "dd_batch_bp");
}

void DDBatchPlan::config(int& fp_lanes, int& bp_lanes, int& fp_persistent, int& bp_persistent, int& fp_streams,
                         int& bp_streams) const {
    fp_lanes = impl_->fp_lpc > 0 ? impl_->fp_lpc : tuning_get("dd_fp_lanes_per_column", 0);
    bp_lanes = impl_->bp_lpc > 0 ? impl_->bp_lpc : tuning_get("dd_bp_lanes_per_pixel", 0);
    fp_persistent = impl_->pf;
    bp_persistent = impl_->pb;
    fp_streams = impl_->fanout_f;
    bp_streams = impl_->fanout_b;
}

int DDBatchPlan::launches(bool forward_dir) const {
    return (int)(forward_dir ? impl_->fwd.size() : impl_->bwd.size());
}

}  // namespace ctp
