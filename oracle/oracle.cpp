// CPU oracle for the projector hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the arithmetic of
// the reference CUDA kernels of wudufan/CTProjector so that the HIP product path
// can be checked against it.  Nothing in the product links or calls this code:
// only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py do.
//
// Every function is a template over the working type T:
//   T = double : the FP64 oracle (the parity reference);
//   T = float  : an FP32 mirror of the reference's own arithmetic, OpenMP
//                parallel over the same index one CUDA thread used (the CPU
//                baseline timed by bench.py).
//
// Citations are file:line in the reference tree, abbreviated
//   cuda/ = src/ct_projector/kernel/projector/cuda/
//
// Behaviour the reference leaves undefined is defined here (and identically in
// the HIP path):
//   * DD BoxInt BP truncates the detector row index and never clamps it
//     (cuda/distanceDriven2DBoxInt.cu:350,383): rows outside [0,nv) read as 0.
//   * InterpolateXZ/YZ never clamp the lower z index (cuda/cudaMath.cu:261-330):
//     it is clamped to [0,nz-1] here.
//   * Parallel FP with bit2 (force FBP weight) reverse-weights batch 0 once per
//     batch (cuda/distanceDriven2DBoxInt.cu:283-287): here every batch item is
//     reverse-weighted once.
//   * Tomo BoxInt BP backprojects batch 0 for every batch
//     (cuda/distanceDriven3DBoxInt.cu:257-259), tomo branchless BP pre-weights
//     view 0 only (cuda/distanceDriven3DBranchless.cu:347) and tomo branchless
//     FP indexes a one-batch accumulator per batch (:209): here each batch item
//     is handled on its own, with every view pre-weighted.
//
// Parity status of each function (what pins it) is listed in tests/ and in the
// repository design notes; see tests/test_oracle_*.py.

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <complex>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// constants exactly as the reference defines them
// ---------------------------------------------------------------------------
const float kPIf = 3.141592657f;        // cuda/projector.h:5
const float kEpsf = 1e-6f;              // cuda/siddon.cu:11
const float kDDMarginf = 0.12f;         // cuda/distanceDriven2DBoxInt.cu:18

template <typename T> struct V2 { T x, y; };
template <typename T> struct V3 { T x, y, z; };

template <typename T>
struct GridT {                           // cuda/projector.h:7-18
    long nx, ny, nz;
    T dx, dy, dz, cx, cy, cz;
};

template <typename T>
struct DetT {                            // cuda/projector.h:20-28
    long nu, nv;
    T du, dv, off_u, off_v;
};

template <typename T> inline T tabs(T v) { return v < 0 ? -v : v; }
template <typename T> inline T tsqrt(T v) { return std::sqrt(v); }
template <typename T> inline T tsin(T v) { return std::sin(v); }
template <typename T> inline T tcos(T v) { return std::cos(v); }
template <typename T> inline T tatan(T v) { return std::atan(v); }

// PhysicsToImg / ImgToPhysics, cuda/projector.h:63-97
template <typename T>
inline V3<T> phys_to_img(V3<T> p, const GridT<T>& g) {
    p.x = (p.x - g.cx) / g.dx + T(g.nx) / T(2);
    p.y = (p.y - g.cy) / g.dy + T(g.ny) / T(2);
    p.z = (p.z - g.cz) / g.dz + T(g.nz) / T(2);
    return p;
}
template <typename T>
inline V2<T> phys_to_img2(V2<T> p, const GridT<T>& g) {
    p.x = (p.x - g.cx) / g.dx + T(g.nx) / T(2);
    p.y = (p.y - g.cy) / g.dy + T(g.ny) / T(2);
    return p;
}
template <typename T>
inline V3<T> img_to_phys(V3<T> p, const GridT<T>& g) {
    p.x = (p.x - T(g.nx) / T(2)) * g.dx + g.cx;
    p.y = (p.y - T(g.ny) / T(2)) * g.dy + g.cy;
    p.z = (p.z - T(g.nz) / T(2)) * g.dz + g.cz;
    return p;
}

// UVToCart, cuda/cudaMath.h:104-118
template <typename T>
inline V3<T> uv_to_cart(T u, T v, const V3<T>& c, const V3<T>& U, const V3<T>& W) {
    return V3<T>{c.x + U.x * u + W.x * v, c.y + U.y * u + W.y * v, c.z + U.z * u + W.z * v};
}

template <typename T>
inline V3<T> load3(const float* p, long i) {
    return V3<T>{T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])};
}

// ClampFloat / Clamp, cuda/cudaMath.cu:163-196
template <typename T>
inline T clampf(T x, T s, T e) { return x < s ? s : (x > e ? e : x); }
inline long clampi(long x, long s, long e) { return x < s ? s : (x >= e ? e - 1 : x); }

// ===========================================================================
// Siddon ray tracing, cuda/siddon.cu
// ===========================================================================
template <typename T>
struct SiddonVars {                      // cuda/siddon.h:7-16
    V3<T> alpha, dAlpha;
    long ngx, ngy, ngz;
    int posx, posy, posz;
    T alphaNow, alphaPrev;
    bool hit;
};

// InboundAlpha, cuda/siddon.cu:13-32
template <typename T>
inline bool inbound_alpha(T& a0, T& aN, T d, T s, long n) {
    const T eps = T(kEpsf);
    if (tabs(d) > eps) {
        a0 = (T(0) - s) / d;
        aN = (T(n) - s) / d;
    } else {
        if (s < 0 || s > T(n)) return false;
    }
    return true;
}

// InboundFirstVoxel, cuda/siddon.cu:35-51  (int conversion truncates)
template <typename T>
inline long inbound_first_voxel(T p, T d, long n) {
    long ng = (long)(int)p;
    if (tabs(d) > T(kEpsf)) {
        if (ng < 0) ng = 0;
        else if (ng >= n) ng = n - 1;
    }
    return ng;
}

// OutboundFirstVoxel, cuda/siddon.cu:54-68
template <typename T>
inline T outbound_first_voxel(long ng, T d) {
    return d > T(kEpsf) ? T(ng + 1) : T(ng);
}

// InitializeSiddon, cuda/siddon.cu:80-149
template <typename T>
SiddonVars<T> initialize_siddon(V3<T> src, V3<T> dst, const GridT<T>& g) {
    SiddonVars<T> v{};
    v.hit = false;
    T dist = tsqrt((src.x - dst.x) * (src.x - dst.x) + (src.y - dst.y) * (src.y - dst.y) +
                   (src.z - dst.z) * (src.z - dst.z));
    src = phys_to_img(src, g);
    dst = phys_to_img(dst, g);
    V3<T> d{dst.x - src.x, dst.y - src.y, dst.z - src.z};
    V3<T> a0{1, 1, 1}, a1{0, 0, 0};
    if (!inbound_alpha(a0.x, a1.x, d.x, src.x, g.nx) || !inbound_alpha(a0.y, a1.y, d.y, src.y, g.ny) ||
        !inbound_alpha(a0.z, a1.z, d.z, src.z, g.nz))
        return v;
    T amin = std::max(T(0), std::max(std::min(a0.x, a1.x), std::max(std::min(a0.y, a1.y), std::min(a0.z, a1.z))));
    T amax = std::min(T(1), std::min(std::max(a0.x, a1.x), std::min(std::max(a0.y, a1.y), std::max(a0.z, a1.z))));
    if (amax <= amin) return v;
    V3<T> p0{src.x + amin * d.x, src.y + amin * d.y, src.z + amin * d.z};
    v.ngx = inbound_first_voxel(p0.x, d.x, g.nx);
    v.ngy = inbound_first_voxel(p0.y, d.y, g.ny);
    v.ngz = inbound_first_voxel(p0.z, d.z, g.nz);
    V3<T> p1{outbound_first_voxel(v.ngx, d.x), outbound_first_voxel(v.ngy, d.y), outbound_first_voxel(v.ngz, d.z)};
    V3<T> al{2, 2, 2}, da{0, 0, 0};
    const T eps = T(kEpsf);
    if (tabs(d.x) > eps) { al.x = (p1.x - src.x) / d.x; da.x = T(1) / tabs(d.x); }
    if (tabs(d.y) > eps) { al.y = (p1.y - src.y) / d.y; da.y = T(1) / tabs(d.y); }
    if (tabs(d.z) > eps) { al.z = (p1.z - src.z) / d.z; da.z = T(1) / tabs(d.z); }
    v.alpha = V3<T>{al.x * dist, al.y * dist, al.z * dist};
    v.dAlpha = V3<T>{da.x * dist, da.y * dist, da.z * dist};
    v.posx = d.x > 0; v.posy = d.y > 0; v.posz = d.z > 0;
    v.alphaNow = v.alphaPrev = amin * dist;
    v.hit = true;
    return v;
}

// SiddonRayTracing (Fwd=true) and SiddonRayTracingTransposeAtomicAdd (Fwd=false),
// cuda/siddon.cu:151-233 and :235-314.  The visitor receives (voxel index, length).
template <typename T, typename Visit>
void siddon_trace(V3<T> src, V3<T> dst, const GridT<T>& g, Visit visit) {
    SiddonVars<T> v = initialize_siddon(src, dst, g);
    if (!v.hit || v.ngx < 0 || v.ngx >= g.nx || v.ngy < 0 || v.ngy >= g.ny || v.ngz < 0 || v.ngz >= g.nz)
        return;
    const long nxy = g.nx * g.ny;
    long idx = v.ngz * nxy + v.ngy * g.nx + v.ngx;
    bool tracing = true;
    while (tracing) {
        long move;
        if (v.alpha.x < v.alpha.y && v.alpha.x < v.alpha.z) {
            v.alphaNow = v.alpha.x;
            v.alpha.x += v.dAlpha.x;
            if (v.posx) { v.ngx++; move = 1; if (v.ngx >= g.nx) tracing = false; }
            else { v.ngx--; move = -1; if (v.ngx < 0) tracing = false; }
        } else if (v.alpha.y < v.alpha.z) {
            v.alphaNow = v.alpha.y;
            v.alpha.y += v.dAlpha.y;
            if (v.posy) { v.ngy++; move = g.nx; if (v.ngy >= g.ny) tracing = false; }
            else { v.ngy--; move = -g.nx; if (v.ngy < 0) tracing = false; }
        } else {
            v.alphaNow = v.alpha.z;
            v.alpha.z += v.dAlpha.z;
            if (v.posz) { v.ngz++; move = nxy; if (v.ngz >= g.nz) tracing = false; }
            else { v.ngz--; move = -nxy; if (v.ngz < 0) tracing = false; }
        }
        visit(idx, v.alphaNow - v.alphaPrev);
        v.alphaPrev = v.alphaNow;
        idx += move;
    }
}

// MoveSourceDstNearGrid, cuda/siddon.cu:318-338 (vecNormalize cuda/cudaMath.h:90-101)
template <typename T>
void move_src_dst_near_grid(V3<T>& s, V3<T>& d, const GridT<T>& g) {
    T ex = T(g.nx) * g.dx, ey = T(g.ny) * g.dy, ez = T(g.nz) * g.dz;
    T r = tsqrt(ex * ex + ey * ey + ez * ez);
    V3<T> w{d.x - s.x, d.y - s.y, d.z - s.z};
    T len = tsqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    V3<T> dir = w;
    if (!(len < T(1e-6f))) dir = V3<T>{w.x / len, w.y / len, w.z / len};
    T t0 = (g.cx - s.x) * dir.x + (g.cy - s.y) * dir.y + (g.cz - s.z) * dir.z;
    T t1 = (d.x - g.cx) * dir.x + (d.y - g.cy) * dir.y + (d.z - g.cz) * dir.z;
    if (t0 > r) s = V3<T>{s.x + (t0 - r) * dir.x, s.y + (t0 - r) * dir.y, s.z + (t0 - r) * dir.z};
    if (t1 > r) d = V3<T>{d.x - (t1 - r) * dir.x, d.y - (t1 - r) * dir.y, d.z - (t1 - r) * dir.z};
}

// Cone-beam ray of detector cell (iu,iv) in view iview, cuda/siddonCone.cu:26-34
template <typename T>
inline void cone_ray(long iu, long iv, long iview, const DetT<T>& det, const float* pc, const float* pu,
                     const float* pvv, const float* ps, const GridT<T>& g, V3<T>& src, V3<T>& dst) {
    T u = (T(iu) - det.off_u - T(det.nu - 1) / T(2)) * det.du;
    T v = (T(iv) - det.off_v - T(det.nv - 1) / T(2)) * det.dv;
    src = load3<T>(ps, iview);
    dst = uv_to_cart(u, v, load3<T>(pc, iview), load3<T>(pu, iview), load3<T>(pvv, iview));
    move_src_dst_near_grid(src, dst, g);
}

template <typename T>
void siddon_cone_fp(T* prj, const float* img, const float* pc, const float* pu, const float* pv, const float* ps,
                    long B, const GridT<T>& g, const DetT<T>& det, long nview) {
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        const float* im = img + ib * nimg;
        T* pr = prj + ib * nprj;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    V3<T> s, d;
                    cone_ray<T>(iu, iv, iview, det, pc, pu, pv, ps, g, s, d);
                    T val = 0;
                    siddon_trace<T>(s, d, g, [&](long idx, T len) { val += T(im[idx]) * len; });
                    pr[(iview * det.nv + iv) * det.nu + iu] += val;
                }
    }
}

template <typename T>
void siddon_cone_bp(T* img, const float* prj, const float* pc, const float* pu, const float* pv, const float* ps,
                    long B, const GridT<T>& g, const DetT<T>& det, long nview) {
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    V3<T> s, d;
                    cone_ray<T>(iu, iv, iview, det, pc, pu, pv, ps, g, s, d);
                    T p = T(pr[(iview * det.nv + iv) * det.nu + iu]);
                    siddon_trace<T>(s, d, g, [&](long idx, T len) {
                        T add = p * len;
#pragma omp atomic
                        im[idx] += add;
                    });
                }
    }
}

// Siddon fan beam, cuda/siddonFan.cu:12-80: source (dso sin t, -dso cos t, z),
// detector point on the arc, both at the row's height z (horizontal rays); no
// MoveSourceDstNearGrid (that is applied to cone beams only).
template <typename T>
void fan_siddon_ray(long iu, long iv, T deg, const DetT<T>& det, T dsd, T dso, V3<T>& s, V3<T>& d) {
    const T c = tcos(deg), sn = tsin(deg);
    const T a = (-(T(iu) - T(det.nu - 1) / T(2)) - det.off_u) * det.du;
    const T z = (T(iv) - T(det.nv - 1) / T(2) + det.off_v) * det.dv;
    s = V3<T>{dso * sn, -dso * c, z};
    const T rx = dsd * tsin(a), ry = -dso + dsd * tcos(a);
    d = V3<T>{rx * c - ry * sn, rx * sn + ry * c, z};
}

template <typename T>
void siddon_fan_fp(T* prj, const float* img, const float* deg, long B, const GridT<T>& g, const DetT<T>& det,
                   long nview, T dsd, T dso) {
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        const float* im = img + ib * nimg;
        T* pr = prj + ib * nprj;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    V3<T> s, d;
                    fan_siddon_ray<T>(iu, iv, T(deg[iview]), det, dsd, dso, s, d);
                    T val = 0;
                    siddon_trace<T>(s, d, g, [&](long idx, T len) { val += T(im[idx]) * len; });
                    pr[(iview * det.nv + iv) * det.nu + iu] += val;
                }
    }
}

template <typename T>
void siddon_fan_bp(T* img, const float* prj, const float* deg, long B, const GridT<T>& g, const DetT<T>& det,
                   long nview, T dsd, T dso) {
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    V3<T> s, d;
                    fan_siddon_ray<T>(iu, iv, T(deg[iview]), det, dsd, dso, s, d);
                    const T p = T(pr[(iview * det.nv + iv) * det.nu + iu]);
                    siddon_trace<T>(s, d, g, [&](long idx, T len) {
                        T add = p * len;
#pragma omp atomic
                        im[idx] += add;
                    });
                }
    }
}

// ===========================================================================
// Box integrals, cuda/cudaMath.cu:336-602
// ===========================================================================
// IntegralBoxX (also used as IntegralBoxY through a stride), cuda/cudaMath.cu:533-565.
// 'row' points at the first element of the line, 'stride' is the element step.
template <typename T, typename S>
T integral_box_1d(const S* row, long stride, T x1, T x2, long n) {
    x1 = clampf(x1, T(0), T(n));
    x2 = clampf(x2, T(0), T(n));
    long i1 = clampi((long)(int)x1, 0, n);
    long i2 = clampi((long)(int)x2, 0, n);
    if (x2 <= x1) return 0;
    T s = 0;
    for (long i = i1; i <= i2; i++) s += T(row[i * stride]);
    s -= T(row[i1 * stride]) * (x1 - T(i1));
    s -= T(row[i2 * stride]) * (T(i2 + 1) - x2);
    return s;
}

// IntegralBoxXY, cuda/cudaMath.cu:336-397; buff is the plane [ny][nx].
template <typename T>
T integral_box_xy(const float* plane, T x1, T y1, T x2, T y2, long nx, long ny) {
    x1 = clampf(x1, T(0), T(nx)); y1 = clampf(y1, T(0), T(ny));
    x2 = clampf(x2, T(0), T(nx)); y2 = clampf(y2, T(0), T(ny));
    long ix1 = clampi((long)(int)x1, 0, nx), iy1 = clampi((long)(int)y1, 0, ny);
    long ix2 = clampi((long)(int)x2, 0, nx), iy2 = clampi((long)(int)y2, 0, ny);
    if (x2 <= x1 || y2 <= y1) return 0;
    auto B = [&](long iy, long ix) { return T(plane[iy * nx + ix]); };
    T s = 0;
    for (long iy = iy1; iy <= iy2; iy++)
        for (long ix = ix1; ix <= ix2; ix++) s += B(iy, ix);
    for (long iy = iy1; iy <= iy2; iy++) s -= B(iy, ix1) * (x1 - T(ix1));
    for (long iy = iy1; iy <= iy2; iy++) s -= B(iy, ix2) * (T(ix2 + 1) - x2);
    for (long ix = ix1; ix <= ix2; ix++) s -= B(iy1, ix) * (y1 - T(iy1));
    for (long ix = ix1; ix <= ix2; ix++) s -= B(iy2, ix) * (T(iy2 + 1) - y2);
    s += B(iy1, ix1) * (x1 - T(ix1)) * (y1 - T(iy1));
    s += B(iy1, ix2) * (T(ix2 + 1) - x2) * (y1 - T(iy1));
    s += B(iy2, ix1) * (x1 - T(ix1)) * (T(iy2 + 1) - y2);
    s += B(iy2, ix2) * (T(ix2 + 1) - x2) * (T(iy2 + 1) - y2);
    return s;
}

// Bilinear interpolation, cuda/cudaMath.cu:199-330.  Generic over the stored
// type; 'a' is the fast axis of length na, 'b' the slow axis of length nb.
// Returns 0 outside [0,na)x[0,nb) when truncate is set.
template <typename T, typename S>
T interp2(const S* plane, long sa, long sb, T a, T b, long na, long nb, bool truncate) {
    if (truncate && (a < 0 || a >= T(na) || b < 0 || b >= T(nb))) return 0;
    a = clampf(a, T(0), T(na));
    b = clampf(b, T(0), T(nb));
    long ia = (long)(int)a, ib = (long)(int)b;
    long ia1 = ia + 1, ib1 = ib + 1;
    ia = clampi(ia, 0, na); ib = clampi(ib, 0, nb);
    ia1 = clampi(ia1, 0, na); ib1 = clampi(ib1, 0, nb);
    T wa = a - T(ia), wb = b - T(ib);
    return T(plane[ib * sb + ia * sa]) * (1 - wa) * (1 - wb) + T(plane[ib * sb + ia1 * sa]) * wa * (1 - wb) +
           T(plane[ib1 * sb + ia * sa]) * (1 - wa) * wb + T(plane[ib1 * sb + ia1 * sa]) * wa * wb;
}

// ===========================================================================
// Distance-driven 2D (fan / parallel), cuda/distanceDriven2D*.cu
// ===========================================================================
struct DDType {                          // cuda/distanceDriven2D.cu:12-25
    int t;
    bool fbp() const { return (t & 1) == 1; }
    bool branchless() const { return (t & 2) == 2; }
    bool force_fbp() const { return (t & 4) == 4; }
};

// Per-view trigonometry.  The reference uses __cosf/__sinf; the FP64 oracle
// evaluates cos/sin in double of the float32 angle.
template <typename T>
inline void view_trig(const float* deg, long iview, T& c, T& s) {
    T d = T(deg[iview]);
    c = tcos(d);
    s = tsin(d);
}

// DDFPTracing (BoxInt), cuda/distanceDriven2DBoxInt.cu:21-106.
// The image accessor im(iz) returns a plane pointer [ny][nx].
template <typename T>
T ddfp_tracing_boxint(const float* img, V2<T> s1, V2<T> s2, V2<T> d1, V2<T> d2, T z, T deg, T a, T cosD, T sinD,
                      const GridT<T>& g) {
    z = clampf(z, T(0), T(g.nz));
    long iz = (long)(int)z;
    long iz1 = iz + 1;
    iz = clampi(iz, 0, g.nz);
    iz1 = clampi(iz1, 0, g.nz);
    T wz = z - T(iz);
    const long nxy = g.nx * g.ny;
    const float* p0 = img + iz * nxy;
    const float* p1 = img + iz1 * nxy;
    const T margin = T(kDDMarginf);
    T valx = 0, valy = 0;
    if (tabs(cosD) > tabs(sinD) - margin) {
        T r1 = (d1.x - s1.x) / (d1.y - s1.y);
        T r2 = (d2.x - s2.x) / (d2.y - s2.y);
        for (long iy = 0; iy < g.ny; iy++) {
            T x1 = s1.x + r1 * (T(iy) + T(0.5) - s1.y);
            T x2 = s2.x + r2 * (T(iy) + T(0.5) - s2.y);
            if (x1 > x2) std::swap(x1, x2);
            valx += (integral_box_1d<T>(p0 + iy * g.nx, 1, x1, x2, g.nx) * (1 - wz) +
                     integral_box_1d<T>(p1 + iy * g.nx, 1, x1, x2, g.nx) * wz) /
                    (x2 - x1);
        }
        valx *= g.dy / tabs(tcos(deg - a));
    }
    if (tabs(cosD) < tabs(sinD) + margin) {
        T r1 = (d1.y - s1.y) / (d1.x - s1.x);
        T r2 = (d2.y - s2.y) / (d2.x - s2.x);
        for (long ix = 0; ix < g.nx; ix++) {
            T y1 = s1.y + r1 * (T(ix) + T(0.5) - s1.x);
            T y2 = s2.y + r2 * (T(ix) + T(0.5) - s2.x);
            if (y1 > y2) std::swap(y1, y2);
            valy += (integral_box_1d<T>(p0 + ix, g.nx, y1, y2, g.ny) * (1 - wz) +
                     integral_box_1d<T>(p1 + ix, g.nx, y1, y2, g.ny) * wz) /
                    (y2 - y1);
        }
        valy *= g.dx / tabs(tsin(deg - a));
    }
    T wx = clampf((tabs(cosD) - tabs(sinD) + margin) / (2 * margin), T(0), T(1));
    return valx * wx + valy * (1 - wx);
}

// Prefix sums for branchless DD, cuda/distanceDriven2DBranchless.cu:17-53.
// accx: [nz][ny][nx+1], accy: [nz][ny+1][nx]; the reference accumulates in float.
template <typename T>
void accumulate_xy(std::vector<T>& accx, std::vector<T>& accy, const float* img, const GridT<T>& g) {
    const long nx = g.nx, ny = g.ny, nz = g.nz;
    accx.assign((nx + 1) * ny * nz, T(0));
    accy.assign(nx * (ny + 1) * nz, T(0));
    for (long iz = 0; iz < nz; iz++) {
        for (long iy = 0; iy < ny; iy++) {
            T* d = accx.data() + iz * (nx + 1) * ny + iy * (nx + 1);
            const float* s = img + iz * nx * ny + iy * nx;
            d[0] = 0;
            for (long ix = 0; ix < nx; ix++) d[ix + 1] = d[ix] + T(s[ix]);
        }
        for (long ix = 0; ix < nx; ix++) {
            T* d = accy.data() + iz * nx * (ny + 1) + ix;
            const float* s = img + iz * nx * ny + ix;
            d[0] = 0;
            for (long iy = 0; iy < ny; iy++) d[(iy + 1) * nx] = d[iy * nx] + T(s[iy * nx]);
        }
    }
}

// InterpolateXZ / InterpolateYZ (non-truncating), cuda/cudaMath.cu:261-330,
// with the lower z index clamped (undefined in the reference when z == nz).
template <typename T>
T interp_line_z(const T* acc, long stride_line, long stride_z, long nline, long nz, T a, T z) {
    a = clampf(a, T(0), T(nline));
    z = clampf(z, T(0), T(nz));
    long ia = (long)(int)a, iz = (long)(int)z;
    long ia1 = ia + 1, iz1 = iz + 1;
    ia = clampi(ia, 0, nline); ia1 = clampi(ia1, 0, nline);
    iz1 = clampi(iz1, 0, nz);
    long izc = clampi(iz, 0, nz);
    T wa = a - T(ia), wz = z - T(iz);
    return acc[izc * stride_z + ia * stride_line] * (1 - wa) * (1 - wz) +
           acc[izc * stride_z + ia1 * stride_line] * wa * (1 - wz) + acc[iz1 * stride_z + ia * stride_line] * (1 - wa) * wz +
           acc[iz1 * stride_z + ia1 * stride_line] * wa * wz;
}

// DDFPTracing (branchless), cuda/distanceDriven2DBranchless.cu:56-129
template <typename T>
T ddfp_tracing_branchless(const std::vector<T>& accx, const std::vector<T>& accy, V2<T> s1, V2<T> s2, V2<T> d1,
                          V2<T> d2, T z, T deg, T a, T cosD, T sinD, const GridT<T>& g) {
    const long nx = g.nx, ny = g.ny, nz = g.nz;
    T val = 0;
    if (tabs(cosD) > tabs(sinD)) {
        T r1 = (d1.x - s1.x) / (d1.y - s1.y);
        T r2 = (d2.x - s2.x) / (d2.y - s2.y);
        for (long iy = 0; iy < ny; iy++) {
            T x1 = s1.x + r1 * (T(iy) + T(0.5) - s1.y);
            T x2 = s2.x + r2 * (T(iy) + T(0.5) - s2.y);
            const T* base = accx.data() + iy * (nx + 1);
            // InterpolateXZ(pAccX, x, iy, z, nx+1, ny, nz)
            T v2 = interp_line_z<T>(base, 1, (nx + 1) * ny, nx + 1, nz, x2, z);
            T v1 = interp_line_z<T>(base, 1, (nx + 1) * ny, nx + 1, nz, x1, z);
            val += (v2 - v1) / (x2 - x1);
        }
        val *= g.dy / tabs(tcos(deg - a));
    } else {
        T r1 = (d1.y - s1.y) / (d1.x - s1.x);
        T r2 = (d2.y - s2.y) / (d2.x - s2.x);
        for (long ix = 0; ix < nx; ix++) {
            T y1 = s1.y + r1 * (T(ix) + T(0.5) - s1.x);
            T y2 = s2.y + r2 * (T(ix) + T(0.5) - s2.x);
            const T* base = accy.data() + ix;
            // InterpolateYZ(pAccY, ix, y, z, nx, ny+1, nz)
            T v2 = interp_line_z<T>(base, nx, nx * (ny + 1), ny + 1, nz, y2, z);
            T v1 = interp_line_z<T>(base, nx, nx * (ny + 1), ny + 1, nz, y1, z);
            val += (v2 - v1) / (y2 - y1);
        }
        val *= g.dx / tabs(tsin(deg - a));
    }
    return val;
}

// Fan ray edges of cell (iu, iv), cuda/distanceDriven2DBoxInt.cu:209-227
template <typename T>
inline void fan_edges(long iu, long iv, T cosD, T sinD, const DetT<T>& det, T dsd, T dso, const GridT<T>& g,
                      V2<T>& src, V2<T>& dst1, V2<T>& dst2, T& a, T& z) {
    a = (-(T(iu) - T(det.nu - 1) / T(2)) - det.off_u) * det.du;
    z = (T(iv) - T(det.nv - 1) / T(2) + det.off_v) * det.dv;
    src = V2<T>{dso * sinD, -dso * cosD};
    V2<T> e1{dsd * tsin(a - det.du / 2), -dso + dsd * tcos(a - det.du / 2)};
    V2<T> e2{dsd * tsin(a + det.du / 2), -dso + dsd * tcos(a + det.du / 2)};
    dst1 = V2<T>{e1.x * cosD - e1.y * sinD, e1.x * sinD + e1.y * cosD};
    dst2 = V2<T>{e2.x * cosD - e2.y * sinD, e2.x * sinD + e2.y * cosD};
    src = phys_to_img2(src, g);
    dst1 = phys_to_img2(dst1, g);
    dst2 = phys_to_img2(dst2, g);
    z = (z - g.cz) / g.dz + T(g.nz) / T(2) - T(0.5);
}

// Parallel ray edges of cell (iu, iv), cuda/distanceDriven2DBoxInt.cu:139-164
template <typename T>
inline void par_edges(long iu, long iv, T cosD, T sinD, const DetT<T>& det, const GridT<T>& g, V2<T>& s1, V2<T>& s2,
                      V2<T>& d1, V2<T>& d2, T& z) {
    T u = (-(T(iu) - T(det.nu - 1) / T(2)) - det.off_u) * det.du;
    z = (T(iv) - T(det.nv - 1) / T(2) + det.off_v) * det.dv;
    T dso = T(g.nx) * g.dx + T(g.ny) * g.dy;
    auto rot = [&](V2<T> p) { return V2<T>{p.x * cosD - p.y * sinD, p.x * sinD + p.y * cosD}; };
    s1 = phys_to_img2(rot(V2<T>{u - det.du / 2, -dso}), g);
    s2 = phys_to_img2(rot(V2<T>{u + det.du / 2, -dso}), g);
    d1 = phys_to_img2(rot(V2<T>{u - det.du / 2, dso}), g);
    d2 = phys_to_img2(rot(V2<T>{u + det.du / 2, dso}), g);
    z = (z - g.cz) / g.dz + T(g.nz) / T(2) - T(0.5);
}

// Fan / parallel DD forward projection: cuda/distanceDriven2DBoxInt.cu:121-299
// and cuda/distanceDriven2DBranchless.cu:145-360.  Output is overwritten.
template <typename T>
void dd2d_fp(T* prj, const float* img, const float* deg, long B, const GridT<T>& g, const DetT<T>& det, long nview,
             T dsd, T dso, int type, bool fan) {
    DDType tp{type};
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        const float* im = img + ib * nimg;
        T* pr = prj + ib * nprj;
        std::vector<T> accx, accy;
        if (tp.branchless()) accumulate_xy<T>(accx, accy, im, g);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    T c, s;
                    view_trig<T>(deg, iview, c, s);
                    T d = T(deg[iview]);
                    T val;
                    if (fan) {
                        V2<T> src, d1, d2;
                        T a, z;
                        fan_edges<T>(iu, iv, c, s, det, dsd, dso, g, src, d1, d2, a, z);
                        if (tp.branchless()) {
                            if (d1.x > d2.x) std::swap(d1, d2);   // :186-192
                            val = ddfp_tracing_branchless<T>(accx, accy, src, src, d1, d2, z, d, a, c, s, g);
                        } else {
                            val = ddfp_tracing_boxint<T>(im, src, src, d1, d2, z, d, a, c, s, g);
                        }
                    } else {
                        V2<T> s1, s2, d1, d2;
                        T z;
                        par_edges<T>(iu, iv, c, s, det, g, s1, s2, d1, d2, z);
                        if (tp.branchless()) {
                            if (d1.x > d2.x) { std::swap(d1, d2); std::swap(s1, s2); }   // :262-271
                            val = ddfp_tracing_branchless<T>(accx, accy, s1, s2, d1, d2, z, d, T(0), c, s, g);
                        } else {
                            val = ddfp_tracing_boxint<T>(im, s1, s2, d1, d2, z, d, T(0), c, s, g);
                            if (tp.force_fbp()) {
                                // ReverseWeightBPParallelKernel, cuda/distanceDriven2DUtils.cu:68-94
                                if (tabs(c) > tabs(s)) val /= g.dy / tabs(c);
                                else val /= g.dx / tabs(s);
                            }
                        }
                    }
                    pr[(iview * det.nv + iv) * det.nu + iu] = val;
                }
    }
}

// Exact transpose of dd2d_fp in the BoxInt mode (no reference counterpart: the
// reference's BP is not the transpose of its FP).  ddfp_visit_boxint walks the
// same lines, rows and box integrals as ddfp_tracing_boxint and hands every
// (pixel, coefficient) pair of the ray's sum to 'visit', so that
// FP = sum coef * img[pixel] and A^T y scatters coef * y into the pixels.
// IntegralBoxX as coefficients: 1 on cells i1..i2, minus (x1 - i1) on i1,
// minus (i2 + 1 - x2) on i2, nothing when x2 <= x1 after clamping.
template <typename T, typename Visit>
void box_1d_coeffs(long base, long stride, T x1, T x2, long n, T scale, Visit visit) {
    x1 = clampf(x1, T(0), T(n));
    x2 = clampf(x2, T(0), T(n));
    long i1 = clampi((long)(int)x1, 0, n);
    long i2 = clampi((long)(int)x2, 0, n);
    if (x2 <= x1) return;
    for (long i = i1; i <= i2; i++) visit(base + i * stride, scale);
    visit(base + i1 * stride, -scale * (x1 - T(i1)));
    visit(base + i2 * stride, -scale * (T(i2 + 1) - x2));
}

template <typename T, typename Visit>
void ddfp_visit_boxint(V2<T> s1, V2<T> s2, V2<T> d1, V2<T> d2, T z, T deg, T a, T cosD, T sinD, const GridT<T>& g,
                       T rayscale, Visit visit) {
    z = clampf(z, T(0), T(g.nz));
    long iz = (long)(int)z;
    long iz1 = iz + 1;
    iz = clampi(iz, 0, g.nz);
    iz1 = clampi(iz1, 0, g.nz);
    T wz = z - T(iz);
    const long nxy = g.nx * g.ny;
    const T margin = T(kDDMarginf);
    T wx = clampf((tabs(cosD) - tabs(sinD) + margin) / (2 * margin), T(0), T(1));
    if (tabs(cosD) > tabs(sinD) - margin) {
        T r1 = (d1.x - s1.x) / (d1.y - s1.y);
        T r2 = (d2.x - s2.x) / (d2.y - s2.y);
        const T f = rayscale * wx * g.dy / tabs(tcos(deg - a));
        for (long iy = 0; iy < g.ny; iy++) {
            T x1 = s1.x + r1 * (T(iy) + T(0.5) - s1.y);
            T x2 = s2.x + r2 * (T(iy) + T(0.5) - s2.y);
            if (x1 > x2) std::swap(x1, x2);
            box_1d_coeffs<T>(iz * nxy + iy * g.nx, 1, x1, x2, g.nx, f * (1 - wz) / (x2 - x1), visit);
            box_1d_coeffs<T>(iz1 * nxy + iy * g.nx, 1, x1, x2, g.nx, f * wz / (x2 - x1), visit);
        }
    }
    if (tabs(cosD) < tabs(sinD) + margin) {
        T r1 = (d1.y - s1.y) / (d1.x - s1.x);
        T r2 = (d2.y - s2.y) / (d2.x - s2.x);
        const T f = rayscale * (1 - wx) * g.dx / tabs(tsin(deg - a));
        for (long ix = 0; ix < g.nx; ix++) {
            T y1 = s1.y + r1 * (T(ix) + T(0.5) - s1.x);
            T y2 = s2.y + r2 * (T(ix) + T(0.5) - s2.x);
            if (y1 > y2) std::swap(y1, y2);
            box_1d_coeffs<T>(iz * nxy + ix, g.nx, y1, y2, g.ny, f * (1 - wz) / (y2 - y1), visit);
            box_1d_coeffs<T>(iz1 * nxy + ix, g.nx, y1, y2, g.ny, f * wz / (y2 - y1), visit);
        }
    }
}

// img += A^T prj for the BoxInt fan / parallel forward projector (type bit 2
// honoured for the parallel beam; the branchless bit is not supported here).
// Deterministic: per-thread partial images summed in thread order.
template <typename T>
void dd2d_fp_adjoint(T* img, const float* prj, const float* deg, long B, const GridT<T>& g, const DetT<T>& det,
                     long nview, T dsd, T dso, int type, bool fan) {
    DDType tp{type};
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
        int nth = 1;
#ifdef _OPENMP
        nth = omp_get_max_threads();
#endif
        std::vector<std::vector<T>> part(nth);
#pragma omp parallel
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            std::vector<T>& acc = part[tid];
            acc.assign(nimg, T(0));
            auto visit = [&acc](long p, T c) { acc[p] += c; };
#pragma omp for collapse(2) schedule(dynamic, 4)
            for (long iview = 0; iview < nview; iview++)
                for (long iv = 0; iv < det.nv; iv++)
                    for (long iu = 0; iu < det.nu; iu++) {
                        const T y = T(pr[(iview * det.nv + iv) * det.nu + iu]);
                        if (y == 0) continue;
                        T c, sn;
                        view_trig<T>(deg, iview, c, sn);
                        T d = T(deg[iview]);
                        if (fan) {
                            V2<T> src, d1, d2;
                            T a, z;
                            fan_edges<T>(iu, iv, c, sn, det, dsd, dso, g, src, d1, d2, a, z);
                            ddfp_visit_boxint<T>(src, src, d1, d2, z, d, a, c, sn, g, y, visit);
                        } else {
                            V2<T> s1, s2, d1, d2;
                            T z;
                            par_edges<T>(iu, iv, c, sn, det, g, s1, s2, d1, d2, z);
                            T ys = y;
                            if (tp.force_fbp()) ys /= (tabs(c) > tabs(sn) ? g.dy / tabs(c) : g.dx / tabs(sn));
                            ddfp_visit_boxint<T>(s1, s2, d1, d2, z, d, T(0), c, sn, g, ys, visit);
                        }
                    }
        }
        for (int t = 0; t < nth; t++)
            if (!part[t].empty())
                for (long k = 0; k < nimg; k++) im[k] += part[t][k];
    }
}

// BP pre-weights, cuda/distanceDriven2DUtils.cu:8-65
template <typename T>
void preweight_bp(std::vector<T>& w, const float* prj, const float* deg, const DetT<T>& det, long nview, T dx, T dy,
                  bool fan) {
    w.resize(det.nu * det.nv * nview);
    for (long iview = 0; iview < nview; iview++) {
        T c, s;
        view_trig<T>(deg, iview, c, s);
        T d = T(deg[iview]);
        for (long iv = 0; iv < det.nv; iv++)
            for (long iu = 0; iu < det.nu; iu++) {
                long k = (iview * det.nv + iv) * det.nu + iu;
                T a = fan ? (-(T(iu) - T(det.nu - 1) / T(2)) - det.off_u) * det.du : T(0);
                T f = tabs(c) > tabs(s) ? dy / tabs(tcos(d - a)) : dx / tabs(tsin(d - a));
                w[k] = T(prj[k]) * f;
            }
    }
}

// GetProjectionOnDetector / ...Parallel / GetFBPWeight, cuda/distanceDriven2DUtils.cu:96-117
template <typename T>
inline T proj_fan(T x, T y, T dso, T c, T s) {
    T rx = x * c + y * s;
    T ry = -x * s + y * c;
    return tatan(rx / (ry + dso));
}
template <typename T>
inline T fbp_weight(T x, T y, T dso, T c, T s) {
    T rx = x * c + y * s;
    T ry = -x * s + y * c;
    return dso * dso / (rx * rx + (dso + ry) * (dso + ry));
}

// Fan / parallel DD backprojection: cuda/distanceDriven2DBoxInt.cu:319-582 and
// cuda/distanceDriven2DBranchless.cu:380-636.  Output accumulates (+=).
template <typename T>
void dd2d_bp(T* img, const float* prj, const float* deg, long B, const GridT<T>& g, const DetT<T>& det, long nview,
             T dsd, T dso, int type, bool fan) {
    (void)dsd;
    DDType tp{type};
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    const bool branchless = tp.branchless();
    const bool fbp = branchless ? tp.fbp() : (tp.fbp() || tp.force_fbp());
    const T margin = T(kDDMarginf);
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
        std::vector<T> w;
        if (!fbp) {
            preweight_bp<T>(w, pr, deg, det, nview, g.dx, g.dy, fan);
        } else {
            w.resize(nprj);
            for (long k = 0; k < nprj; k++) w[k] = T(pr[k]);
        }
        std::vector<T> acc;   // branchless: [nview][nv][nu+1]
        if (branchless) {
            acc.assign((det.nu + 1) * det.nv * nview, T(0));
            for (long r = 0; r < det.nv * nview; r++) {
                T* d = acc.data() + r * (det.nu + 1);
                const T* s = w.data() + r * det.nu;
                d[0] = 0;
                for (long iu = 0; iu < det.nu; iu++) d[iu + 1] = d[iu] + s[iu];
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
        for (long iz = 0; iz < g.nz; iz++)
            for (long iy = 0; iy < g.ny; iy++)
                for (long ix = 0; ix < g.nx; ix++) {
                    V3<T> dst = img_to_phys(V3<T>{T(ix), T(iy), T(iz)}, g);
                    dst.x += g.dx / 2; dst.y += g.dy / 2; dst.z += g.dz / 2;
                    T x1 = dst.x - g.dx / T(2), x2 = dst.x + g.dx / T(2);
                    T y1 = dst.y - g.dy / T(2), y2 = dst.y + g.dy / T(2);
                    T ivf = dst.z / det.dv + det.off_v + T(det.nv) / T(2) - T(0.5);
                    long ivi = (long)(int)ivf;   // BoxInt passes the float row into an int parameter
                    T sum = 0;
                    for (long iview = 0; iview < nview; iview++) {
                        T c, s;
                        view_trig<T>(deg, iview, c, s);
                        // row/plane of this view in the weighted projections
                        auto box = [&](T u1, T u2) -> T {
                            if (u1 > u2) std::swap(u1, u2);
                            if (branchless) {
                                const T* plane = acc.data() + iview * det.nv * (det.nu + 1);
                                T v2 = interp2<T, T>(plane, 1, det.nu + 1, u2, ivf, det.nu + 1, det.nv, false);
                                T v1 = interp2<T, T>(plane, 1, det.nu + 1, u1, ivf, det.nu + 1, det.nv, false);
                                return (v2 - v1) / (u2 - u1);
                            }
                            if (ivi < 0 || ivi >= det.nv) return T(0);
                            const T* row = w.data() + (iview * det.nv + ivi) * det.nu;
                            return integral_box_1d<T, T>(row, 1, u1, u2, det.nu) / (u2 - u1);
                        };
                        T val;
                        if (fan) {
                            T a1, a2;
                            if (tabs(c) > tabs(s)) {
                                a1 = proj_fan(x1, dst.y, dso, c, s);
                                a2 = proj_fan(x2, dst.y, dso, c, s);
                            } else {
                                a1 = proj_fan(dst.x, y1, dso, c, s);
                                a2 = proj_fan(dst.x, y2, dso, c, s);
                            }
                            a1 = -(a1 / det.du + det.off_u) + T(det.nu) / T(2);
                            a2 = -(a2 / det.du + det.off_u) + T(det.nu) / T(2);
                            val = box(a1, a2);
                            if (fbp) val *= fbp_weight(dst.x, dst.y, dso, c, s);
                        } else if (branchless) {
                            T u1, u2;
                            if (tabs(c) > tabs(s)) { u1 = x1 * c + dst.y * s; u2 = x2 * c + dst.y * s; }
                            else { u1 = dst.x * c + y1 * s; u2 = dst.x * c + y2 * s; }
                            u1 = -(u1 / det.du + det.off_u) + T(det.nu) / T(2);
                            u2 = -(u2 / det.du + det.off_u) + T(det.nu) / T(2);
                            val = box(u1, u2);
                        } else {
                            T valx = 0, valy = 0;
                            if (tabs(c) > tabs(s) - margin) {
                                T u1 = x1 * c + dst.y * s, u2 = x2 * c + dst.y * s;
                                u1 = -(u1 / det.du + det.off_u) + T(det.nu) / T(2);
                                u2 = -(u2 / det.du + det.off_u) + T(det.nu) / T(2);
                                valx = box(u1, u2);
                            }
                            if (tabs(c) < tabs(s) + margin) {
                                T u1 = dst.x * c + y1 * s, u2 = dst.x * c + y2 * s;
                                u1 = -(u1 / det.du + det.off_u) + T(det.nu) / T(2);
                                u2 = -(u2 / det.du + det.off_u) + T(det.nu) / T(2);
                                valy = box(u1, u2);
                            }
                            T wx = clampf((tabs(c) - tabs(s) + margin) / (2 * margin), T(0), T(1));
                            val = valx * wx + valy * (1 - wx);
                        }
                        sum += val;
                    }
                    im[(iz * g.ny + iy) * g.nx + ix] += sum;
                }
    }
}

// ===========================================================================
// Distance-driven tomosynthesis, cuda/distanceDriven3D*.cu
// ===========================================================================
// ProjectConeToDetCart, cuda/distanceDriven3DUtils.cu:45-56
template <typename T>
inline V2<T> project_cone_cart(V3<T> pt, V3<T> c, V3<T> s, const DetT<T>& det) {
    T r = (c.z - s.z) / (pt.z - s.z);
    T u = s.x - c.x + r * (pt.x - s.x);
    T v = s.y - c.y + r * (pt.y - s.y);
    u = u / det.du + det.off_u + T(det.nu - 1) / T(2);
    v = v / det.dv + det.off_v + T(det.nv - 1) / T(2);
    return V2<T>{u, v};
}

// Tomo FP, BoxInt (cuda/distanceDriven3DBoxInt.cu:28-117) and branchless
// (cuda/distanceDriven3DBranchless.cu:78-171).  Detector axes are (1,0,0),(0,1,0)
// (cuda/distanceDriven3D.cu:59-75).  Output overwritten.
template <typename T>
void tomo_fp(T* prj, const float* img, const float* pc, const float* ps, long B, const GridT<T>& g,
             const DetT<T>& det, long nview, int type) {
    const long nx = g.nx, ny = g.ny, nz = g.nz;
    const long nimg = nx * ny * nz, nprj = det.nu * det.nv * nview;
    const V3<T> U{1, 0, 0}, W{0, 1, 0};
    for (long ib = 0; ib < B; ib++) {
        const float* im = img + ib * nimg;
        T* pr = prj + ib * nprj;
        std::vector<double> sat;   // [nz][ny+1][nx+1], always double in the reference
        if (type == 1) {
            sat.assign((nx + 1) * (ny + 1) * nz, 0.0);
            for (long iz = 0; iz < nz; iz++) {
                double* p = sat.data() + iz * (nx + 1) * (ny + 1);
                for (long iy = 0; iy < ny; iy++) {
                    double* d = p + (iy + 1) * (nx + 1);
                    const float* s = im + iz * nx * ny + iy * nx;
                    d[0] = 0;
                    for (long ix = 0; ix < nx; ix++) d[ix + 1] = d[ix] + s[ix];
                }
                for (long ix = 0; ix < nx; ix++) {
                    double* a = p + ix + 1;
                    for (long iy = 0; iy < ny; iy++) a[(iy + 1) * (nx + 1)] += a[iy * (nx + 1)];
                }
            }
        }
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (long iview = 0; iview < nview; iview++)
            for (long iv = 0; iv < det.nv; iv++)
                for (long iu = 0; iu < det.nu; iu++) {
                    T u = (T(iu) - det.off_u - T(det.nu - 1) / T(2)) * det.du;
                    T v = (T(iv) - det.off_v - T(det.nv - 1) / T(2)) * det.dv;
                    V3<T> c = load3<T>(pc, iview);
                    V3<T> sx1 = phys_to_img(uv_to_cart(u - det.du / 2, v, c, U, W), g);
                    V3<T> sx2 = phys_to_img(uv_to_cart(u + det.du / 2, v, c, U, W), g);
                    V3<T> sy1 = phys_to_img(uv_to_cart(u, v - det.dv / 2, c, U, W), g);
                    V3<T> sy2 = phys_to_img(uv_to_cart(u, v + det.dv / 2, c, U, W), g);
                    V3<T> src = phys_to_img(load3<T>(ps, iview), g);
                    if (sx1.x > sx2.x) std::swap(sx1, sx2);
                    if (sy1.y > sy2.y) std::swap(sy1, sy2);
                    T rx1 = (sx1.x - src.x) / (sx1.z - src.z), rx2 = (sx2.x - src.x) / (sx2.z - src.z);
                    T ry1 = (sy1.y - src.y) / (sy1.z - src.z), ry2 = (sy2.y - src.y) / (sy2.z - src.z);
                    T val = 0;
                    for (long iz = 0; iz < nz; iz++) {
                        T x1 = src.x + rx1 * (T(iz) - src.z), x2 = src.x + rx2 * (T(iz) - src.z);
                        T y1 = src.y + ry1 * (T(iz) - src.z), y2 = src.y + ry2 * (T(iz) - src.z);
                        if (type == 1) {
                            const double* p = sat.data() + iz * (nx + 1) * (ny + 1);
                            T s = interp2<T, double>(p, 1, nx + 1, x2, y2, nx + 1, ny + 1, false) +
                                  interp2<T, double>(p, 1, nx + 1, x1, y1, nx + 1, ny + 1, false) -
                                  interp2<T, double>(p, 1, nx + 1, x2, y1, nx + 1, ny + 1, false) -
                                  interp2<T, double>(p, 1, nx + 1, x1, y2, nx + 1, ny + 1, false);
                            val += s / ((x2 - x1) * (y2 - y1));
                        } else {
                            val += integral_box_xy<T>(im + iz * nx * ny, x1, y1, x2, y2, nx, ny) / ((x2 - x1) * (y2 - y1));
                        }
                    }
                    V3<T> dst = uv_to_cart(u, v, c, U, W);
                    V3<T> s0 = load3<T>(ps, iview);
                    val *= g.dz / tabs(s0.z - dst.z) *
                           tsqrt((s0.z - dst.z) * (s0.z - dst.z) + (s0.y - dst.y) * (s0.y - dst.y) +
                                 (s0.x - dst.x) * (s0.x - dst.x));
                    pr[(iview * det.nv + iv) * det.nu + iu] = val;
                }
    }
}

// Tomo BP, BoxInt (cuda/distanceDriven3DBoxInt.cu:170-280) and branchless
// (cuda/distanceDriven3DBranchless.cu:237-404); pre-weight
// PreweightBPCartKernelXY cuda/distanceDriven3DUtils.cu:10-40.  Output overwritten.
template <typename T>
void tomo_bp(T* img, const float* prj, const float* pc, const float* ps, long B, const GridT<T>& g,
             const DetT<T>& det, long nview, int type) {
    const long nx = g.nx, ny = g.ny, nz = g.nz, nu = det.nu, nv = det.nv;
    const long nimg = nx * ny * nz, nprj = nu * nv * nview;
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
        std::vector<T> w(nprj);
        std::vector<float> wf(nprj);
        for (long iview = 0; iview < nview; iview++) {
            V3<T> c = load3<T>(pc, iview), s = load3<T>(ps, iview);
            for (long iv = 0; iv < nv; iv++)
                for (long iu = 0; iu < nu; iu++) {
                    T u = (T(iu) - det.off_u - T(nu - 1) / T(2)) * det.du;
                    T v = (T(iv) - det.off_v - T(nv - 1) / T(2)) * det.dv;
                    V3<T> d = uv_to_cart(u, v, c, V3<T>{1, 0, 0}, V3<T>{0, 1, 0});
                    long k = (iview * nv + iv) * nu + iu;
                    w[k] = T(pr[k]) * (g.dz / tabs(s.z - d.z) *
                                       tsqrt((s.z - d.z) * (s.z - d.z) + (s.y - d.y) * (s.y - d.y) + (s.x - d.x) * (s.x - d.x)));
                    wf[k] = float(w[k]);
                }
        }
        std::vector<double> sat;   // [nview][nv+1][nu+1]
        if (type == 1) {
            sat.assign((nu + 1) * (nv + 1) * nview, 0.0);
            for (long iview = 0; iview < nview; iview++) {
                double* p = sat.data() + iview * (nu + 1) * (nv + 1);
                for (long iv = 0; iv < nv; iv++) {
                    double* d = p + (iv + 1) * (nu + 1);
                    // the reference accumulates the float pre-weighted projections
                    const float* s = wf.data() + (iview * nv + iv) * nu;
                    const T* sd = w.data() + (iview * nv + iv) * nu;
                    d[0] = 0;
                    for (long iu = 0; iu < nu; iu++) d[iu + 1] = d[iu] + (sizeof(T) == 8 ? double(sd[iu]) : double(s[iu]));
                }
                for (long iu = 0; iu < nu; iu++) {
                    double* a = p + iu + 1;
                    for (long iv = 0; iv < nv; iv++) a[(iv + 1) * (nu + 1)] += a[iv * (nu + 1)];
                }
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
        for (long iz = 0; iz < nz; iz++)
            for (long iy = 0; iy < ny; iy++)
                for (long ix = 0; ix < nx; ix++) {
                    T x = (T(ix) - T(nx - 1) / T(2)) * g.dx + g.cx;
                    T y = (T(iy) - T(ny - 1) / T(2)) * g.dy + g.cy;
                    T z = (T(iz) - T(nz - 1) / T(2)) * g.dz + g.cz;
                    T val = 0;
                    for (long iview = 0; iview < nview; iview++) {
                        V3<T> s = load3<T>(ps, iview), c = load3<T>(pc, iview);
                        T u1 = project_cone_cart(V3<T>{x - g.dx / 2, y, z}, c, s, det).x;
                        T u2 = project_cone_cart(V3<T>{x + g.dx / 2, y, z}, c, s, det).x;
                        T v1 = project_cone_cart(V3<T>{x, y - g.dy / 2, z}, c, s, det).y;
                        T v2 = project_cone_cart(V3<T>{x, y + g.dy / 2, z}, c, s, det).y;
                        if (type == 1) {
                            const double* p = sat.data() + iview * (nu + 1) * (nv + 1);
                            T sm = interp2<T, double>(p, 1, nu + 1, u2, v2, nu + 1, nv + 1, false) -
                                   interp2<T, double>(p, 1, nu + 1, u2, v1, nu + 1, nv + 1, false) +
                                   interp2<T, double>(p, 1, nu + 1, u1, v1, nu + 1, nv + 1, false) -
                                   interp2<T, double>(p, 1, nu + 1, u1, v2, nu + 1, nv + 1, false);
                            val += sm / ((u2 - u1) * (v2 - v1));
                        } else {
                            // IntegralBoxXY over the weighted projections of this view
                            T x1 = clampf(u1, T(0), T(nu)), y1c = clampf(v1, T(0), T(nv));
                            T x2 = clampf(u2, T(0), T(nu)), y2c = clampf(v2, T(0), T(nv));
                            T sm = 0;
                            if (x2 > x1 && y2c > y1c) {
                                long ix1 = clampi((long)(int)x1, 0, nu), iy1 = clampi((long)(int)y1c, 0, nv);
                                long ix2 = clampi((long)(int)x2, 0, nu), iy2 = clampi((long)(int)y2c, 0, nv);
                                const T* P = w.data() + iview * nv * nu;
                                auto Bv = [&](long r, long q) { return P[r * nu + q]; };
                                for (long r = iy1; r <= iy2; r++)
                                    for (long q = ix1; q <= ix2; q++) sm += Bv(r, q);
                                for (long r = iy1; r <= iy2; r++) sm -= Bv(r, ix1) * (x1 - T(ix1));
                                for (long r = iy1; r <= iy2; r++) sm -= Bv(r, ix2) * (T(ix2 + 1) - x2);
                                for (long q = ix1; q <= ix2; q++) sm -= Bv(iy1, q) * (y1c - T(iy1));
                                for (long q = ix1; q <= ix2; q++) sm -= Bv(iy2, q) * (T(iy2 + 1) - y2c);
                                sm += Bv(iy1, ix1) * (x1 - T(ix1)) * (y1c - T(iy1));
                                sm += Bv(iy1, ix2) * (T(ix2 + 1) - x2) * (y1c - T(iy1));
                                sm += Bv(iy2, ix1) * (x1 - T(ix1)) * (T(iy2 + 1) - y2c);
                                sm += Bv(iy2, ix2) * (T(ix2 + 1) - x2) * (T(iy2 + 1) - y2c);
                            }
                            val += sm / ((u2 - u1) * (v2 - v1));
                        }
                    }
                    im[(iz * ny + iy) * nx + ix] = val;
                }
    }
}

// ===========================================================================
// Ramp filter and fan FBP backprojection, cuda/fbpMath.cu, cuda/fbp.cu,
// cuda/fbpParallel.cu
// ===========================================================================
// Windowed ramp kernel h_w[i], i in [0, 2nu-1), as produced by GetRamp
// (cuda/fbpMath.cu:90-268) in exact arithmetic: the spatial RL kernel is
// multiplied in the frequency domain by the window and transformed back.  For
// Hamming/Hann that is a 3-tap circular smoothing of the spatial kernel; for
// the cosine window an explicit DFT in double is used.
void ramp_kernel(double* h, long nu, double da, int filter, bool equispace) {
    const long L = 2 * nu - 1;
    const double PI = double(kPIf);
    std::vector<double> r(L);
    for (long i = 0; i < L; i++) {
        long k = i - (nu - 1);
        if (k == 0) r[i] = 1.0 / (4 * da * da);
        else if (k % 2 != 0) {
            if (equispace) r[i] = -1.0 / (PI * PI * double(k) * double(k) * da * da);
            else { double sk = std::sin(double(k) * da); r[i] = -1.0 / (PI * PI * sk * sk); }
        } else r[i] = 0;
    }
    if (filter == 1 || filter == 2) {
        // window w[f] = A + (1-A) cos(2 pi f / L)  <=>  circular taps {A, (1-A)/2, (1-A)/2}
        double A = filter == 1 ? double(0.54f) : double(0.5f);
        double Bt = filter == 1 ? double(0.46f) : double(0.5f);
        for (long i = 0; i < L; i++)
            h[i] = A * r[i] + 0.5 * Bt * (r[(i - 1 + L) % L] + r[(i + 1) % L]);
    } else if (filter == 3) {
        std::vector<std::complex<double>> F(L);
        for (long f = 0; f < L; f++) {
            std::complex<double> s = 0;
            for (long i = 0; i < L; i++) s += r[i] * std::polar(1.0, -2 * M_PI * double((f * i) % L) / double(L));
            F[f] = s * std::abs(std::cos(PI * double(f) / double(L)));
        }
        for (long i = 0; i < L; i++) {
            std::complex<double> s = 0;
            for (long f = 0; f < L; f++) s += F[f] * std::polar(1.0, 2 * M_PI * double((f * i) % L) / double(L));
            h[i] = s.real() / double(L);
        }
    } else {
        for (long i = 0; i < L; i++) h[i] = r[i];
    }
}

// Fan (equispace=false) / parallel (true) ramp filtering, cuda/fbp.cu:15-162 and
// cuda/fbpParallel.cu:14-109, as the exact linear convolution the FFT computes.
template <typename T>
void ramp_filter(T* out, const float* prj, long B, long nu, long nv, long nview, T da, T off_u, T dsd, T dso,
                 int filter, bool equispace) {
    const long L = 2 * nu - 1;
    std::vector<double> hd(L);
    ramp_kernel(hd.data(), nu, double(da), filter, equispace);
    std::vector<T> h(L), pw(nu);
    for (long i = 0; i < L; i++) h[i] = T(hd[i]);
    for (long i = 0; i < nu; i++) {
        T ang = ((T(i) - T(nu - 1) / T(2)) - off_u) * da;
        pw[i] = equispace ? T(1) : tcos(ang);
    }
    const T PI = T(kPIf);
    T scale = equispace ? PI / T(nview) * da : PI / T(nview) * da / dso;
    (void)dsd;
#pragma omp parallel for collapse(2) schedule(static)
    for (long ib = 0; ib < B; ib++)
        for (long r = 0; r < nview * nv; r++) {
            const float* x = prj + ib * nview * nv * nu + r * nu;
            T* y = out + ib * nview * nv * nu + r * nu;
            std::vector<T> xw(nu);
            for (long m = 0; m < nu; m++) xw[m] = T(x[m]) * pw[m];
            for (long j = 0; j < nu; j++) {
                T s = 0;
                for (long m = 0; m < nu; m++) s += xw[m] * h[nu - 1 + j - m];
                y[j] = s * scale;
            }
        }
}

// Fan FBP backprojection bpFanKernel3D, cuda/fbp.cu:264-359 (5 slices per
// thread, bilinear interpolation with truncation).  Output accumulates (+=).
template <typename T>
void fan_fbp_bp(T* img, const float* prj, const float* deg, long B, const GridT<T>& g, const DetT<T>& det, long nview,
                T dsd, T dso, bool fan = true) {
    (void)dsd;
    const long nzb = 5;
    const long nimg = g.nx * g.ny * g.nz, nprj = det.nu * det.nv * nview;
    const long nbatch_z = (g.nz + nzb - 1) / nzb;
    for (long ib = 0; ib < B; ib++) {
        T* im = img + ib * nimg;
        const float* pr = prj + ib * nprj;
#pragma omp parallel for collapse(2) schedule(static)
        for (long izb = 0; izb < nbatch_z; izb++)
            for (long iy = 0; iy < g.ny; iy++)
                for (long ix = 0; ix < g.nx; ix++) {
                    V3<T> pt = img_to_phys(V3<T>{T(ix) + T(0.5), T(iy) + T(0.5), T(izb * nzb) + T(0.5)}, g);
                    T val[nzb] = {0, 0, 0, 0, 0};
                    for (long iview = 0; iview < nview; iview++) {
                        T c, s;
                        view_trig<T>(deg, iview, c, s);
                        T rx = pt.x * c + pt.y * s;
                        T ry = -pt.x * s + pt.y * c;
                        // fan: bpFanKernel3D (cuda/fbp.cu:300-330); parallel:
                        // bpParallelKernel3D (cuda/fbpParallel.cu:224-240), no weight
                        T a = fan ? tatan(rx / (ry + dso)) : rx;
                        T dist = fan ? dso * dso / (rx * rx + (dso + ry) * (dso + ry)) : T(1);
                        T pu = -(a / det.du + det.off_u) + (T(det.nu) - T(1)) / T(2);
                        for (long k = 0; k < nzb; k++) {
                            T pv = (pt.z + T(k) * g.dz) / det.dv + det.off_v + (T(det.nv) - T(1)) / T(2);
                            val[k] += interp2<T, float>(pr + iview * det.nu * det.nv, 1, det.nu, pu, pv, det.nu, det.nv, true) * dist;
                        }
                    }
                    for (long k = 0; k < nzb; k++) {
                        long iz = izb * nzb + k;
                        if (iz < g.nz) im[(iz * g.ny + iy) * g.nx + ix] += val[k];
                    }
                }
    }
}

template <typename T>
GridT<T> mkgrid(long nx, long ny, long nz, float dx, float dy, float dz, float cx, float cy, float cz) {
    return GridT<T>{nx, ny, nz, T(dx), T(dy), T(dz), T(cx), T(cy), T(cz)};
}
template <typename T>
DetT<T> mkdet(long nu, long nv, float du, float dv, float ou, float ov) {
    return DetT<T>{nu, nv, T(du), T(dv), T(ou), T(ov)};
}

}  // namespace

// ===========================================================================
// C ABI of the oracle (suffix _d: FP64 oracle, _f: FP32 mirror)
// ===========================================================================
#define GRIDARGS size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz
#define DETARGS size_t nu, size_t nv, size_t nview, float du, float dv, float off_u, float off_v
#define MKG(T) mkgrid<T>((long)nx, (long)ny, (long)nz, dx, dy, dz, cx, cy, cz)
#define MKD(T) mkdet<T>((long)nu, (long)nv, du, dv, off_u, off_v)

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

#define DEF_SIDDON(SUF, T)                                                                                   \
    void orc_siddon_cone_fp_##SUF(T* prj, const float* img, const float* dc, const float* du_, const float* dv_, \
                                  const float* src, size_t B, GRIDARGS, DETARGS) {                           \
        siddon_cone_fp<T>(prj, img, dc, du_, dv_, src, (long)B, MKG(T), MKD(T), (long)nview);                 \
    }                                                                                                        \
    void orc_siddon_cone_bp_##SUF(T* img, const float* prj, const float* dc, const float* du_, const float* dv_, \
                                  const float* src, size_t B, GRIDARGS, DETARGS) {                           \
        siddon_cone_bp<T>(img, prj, dc, du_, dv_, src, (long)B, MKG(T), MKD(T), (long)nview);                 \
    }
DEF_SIDDON(d, double)
DEF_SIDDON(f, float)

#define DEF_SIDDON_FAN(SUF, T)                                                                               \
    void orc_siddon_fan_fp_##SUF(T* prj, const float* img, const float* deg, size_t B, GRIDARGS, DETARGS,    \
                                 float dsd, float dso) {                                                     \
        siddon_fan_fp<T>(prj, img, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso));                \
    }                                                                                                        \
    void orc_siddon_fan_bp_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS,    \
                                 float dsd, float dso) {                                                     \
        siddon_fan_bp<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso));                \
    }
DEF_SIDDON_FAN(d, double)
DEF_SIDDON_FAN(f, float)

#define DEF_DD2D(SUF, T)                                                                                     \
    void orc_dd_fan_fp_##SUF(T* prj, const float* img, const float* deg, size_t B, GRIDARGS, DETARGS, float dsd, \
                             float dso, int type) {                                                          \
        dd2d_fp<T>(prj, img, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso), type, true);          \
    }                                                                                                        \
    void orc_dd_fan_bp_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS, float dsd, \
                             float dso, int type) {                                                          \
        dd2d_bp<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso), type, true);          \
    }                                                                                                        \
    void orc_dd_par_fp_##SUF(T* prj, const float* img, const float* deg, size_t B, GRIDARGS, DETARGS, int type) { \
        dd2d_fp<T>(prj, img, deg, (long)B, MKG(T), MKD(T), (long)nview, T(1000), T(500), type, false);        \
    }                                                                                                        \
    void orc_dd_par_bp_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS, int type) { \
        dd2d_bp<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(1000), T(500), type, false);        \
    }                                                                                                        \
    void orc_dd_fan_fp_adjoint_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS,   \
                                     float dsd, float dso, int type) {                                       \
        dd2d_fp_adjoint<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso), type, true);  \
    }                                                                                                        \
    void orc_dd_par_fp_adjoint_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS,   \
                                     int type) {                                                             \
        dd2d_fp_adjoint<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(1000), T(500), type, false); \
    }
DEF_DD2D(d, double)
DEF_DD2D(f, float)

#define DEF_TOMO(SUF, T)                                                                                     \
    void orc_dd_tomo_fp_##SUF(T* prj, const float* img, const float* dc, const float* src, size_t B, GRIDARGS,  \
                              DETARGS, int type) {                                                           \
        tomo_fp<T>(prj, img, dc, src, (long)B, MKG(T), MKD(T), (long)nview, type);                            \
    }                                                                                                        \
    void orc_dd_tomo_bp_##SUF(T* img, const float* prj, const float* dc, const float* src, size_t B, GRIDARGS,  \
                              DETARGS, int type) {                                                           \
        tomo_bp<T>(img, prj, dc, src, (long)B, MKG(T), MKD(T), (long)nview, type);                            \
    }
DEF_TOMO(d, double)
DEF_TOMO(f, float)

void orc_ramp_kernel(double* h, size_t nu, double da, int filter, int equispace) {
    ramp_kernel(h, (long)nu, da, filter, equispace != 0);
}

#define DEF_FBP(SUF, T)                                                                                      \
    void orc_fan_filter_##SUF(T* out, const float* prj, size_t B, size_t nu, size_t nv, size_t nview, float da, \
                              float dv, float off_a, float off_v, float dsd, float dso, int filter) {        \
        (void)dv; (void)off_v;                                                                               \
        ramp_filter<T>(out, prj, (long)B, (long)nu, (long)nv, (long)nview, T(da), T(off_a), T(dsd), T(dso), filter, \
                       false);                                                                               \
    }                                                                                                        \
    void orc_par_filter_##SUF(T* out, const float* prj, size_t B, size_t nu, size_t nv, size_t nview, float du, \
                              float dv, float off_u, float off_v, int filter) {                              \
        (void)dv; (void)off_v;                                                                               \
        ramp_filter<T>(out, prj, (long)B, (long)nu, (long)nv, (long)nview, T(du), T(off_u), T(0), T(1), filter, true); \
    }                                                                                                        \
    void orc_fan_fbp_bp_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS, float dsd, \
                              float dso) {                                                                   \
        fan_fbp_bp<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(dsd), T(dso));                  \
    }
DEF_FBP(d, double)
DEF_FBP(f, float)

#define DEF_PARBP(SUF, T)                                                                                    \
    void orc_par_pixel_bp_##SUF(T* img, const float* prj, const float* deg, size_t B, GRIDARGS, DETARGS) {    \
        fan_fbp_bp<T>(img, prj, deg, (long)B, MKG(T), MKD(T), (long)nview, T(1000), T(500), false);           \
    }
DEF_PARBP(d, double)
DEF_PARBP(f, float)

}  // extern "C"
