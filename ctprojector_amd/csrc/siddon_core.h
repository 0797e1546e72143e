// Siddon ray traversal shared by the cone-beam and fan-beam projectors
// (the reference's cuda/siddon.cu): ray set-up and one traversal step.
#pragma once

#include "common.h"

namespace ctp {

// one geometry row (x, y, z) of a read-only [nview][3] table; wave-uniform
// rows are read with scalar loads
__device__ inline float3 ld3(const float* p, int i) {
    return make_float3(const_load(p, 3L * i), const_load(p, 3L * i + 1), const_load(p, 3L * i + 2));
}

// Traversal state of one ray: the reference's InitializeSiddon
// (cuda/siddon.cu:80-149) after MoveSourceDstNearGrid (cuda/siddon.cu:318-338),
// restated with integer crossing counters for the exit test.
struct SiddonRay {
    float ax, ay, az;      // next crossing along each axis (mm from the moved source)
    float dax, day, daz;   // spacing of the crossings along each axis
    float a;               // position of the last crossing (mm)
    int off;               // byte offset of the current voxel
    int mx, my, mz;        // byte offset steps along x, y, z
    int rx, ry, rz;        // crossings left along each axis before the index leaves the grid
};

// s: source, t: detector point (physical).  With sw the x and y axes of the
// grid and of the points are exchanged after the move (volume read as
// [z][x][y]).  move: the cone projector slides both points towards the grid
// first; the fan projector does not (cuda/siddonFan.cu:41-46).  Returns false
// when the ray misses the grid.
// MoveSourceDstNearGrid (cuda/siddon.cu:318-338): slides the source and the
// detector point along their line to within the grid's diagonal of its centre.
__device__ inline void siddon_move(float3& s, float3& t, const Grid& g) {
    const float ex = (float)g.nx * g.dx, ey = (float)g.ny * g.dy, ez = (float)g.nz * g.dz;
    const float rr = sqrtf(ex * ex + ey * ey + ez * ez);
    float3 w = make_float3(t.x - s.x, t.y - s.y, t.z - s.z);
    const float len = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    if (!(len < 1e-6f)) w = make_float3(w.x / len, w.y / len, w.z / len);
    const float t0 = (g.cx - s.x) * w.x + (g.cy - s.y) * w.y + (g.cz - s.z) * w.z;
    const float t1 = (t.x - g.cx) * w.x + (t.y - g.cy) * w.y + (t.z - g.cz) * w.z;
    if (t0 > rr) s = make_float3(s.x + (t0 - rr) * w.x, s.y + (t0 - rr) * w.y, s.z + (t0 - rr) * w.z);
    if (t1 > rr) t = make_float3(t.x - (t1 - rr) * w.x, t.y - (t1 - rr) * w.y, t.z - (t1 - rr) * w.z);
}

// Index-space direction of the (moved) segment from s to t, as the traversal
// set-up computes it (cuda/siddon.cu:86-96).
__device__ inline float3 siddon_index_dir(float3 s, float3 t, const Grid& g) {
    const float3 si = make_float3((s.x - g.cx) / g.dx + g.nx / 2.0f, (s.y - g.cy) / g.dy + g.ny / 2.0f,
                                  (s.z - g.cz) / g.dz + g.nz / 2.0f);
    const float3 ti = make_float3((t.x - g.cx) / g.dx + g.nx / 2.0f, (t.y - g.cy) / g.dy + g.ny / 2.0f,
                                  (t.z - g.cz) / g.dz + g.nz / 2.0f);
    return make_float3(ti.x - si.x, ti.y - si.y, ti.z - si.z);
}

__device__ inline bool siddon_setup(SiddonRay& r, float3 s, float3 t, Grid& g, bool sw, bool move = true) {
    if (move) siddon_move(s, t, g);
    if (sw) {
        const float a = s.x; s.x = s.y; s.y = a;
        const float b = t.x; t.x = t.y; t.y = b;
        const int n = g.nx; g.nx = g.ny; g.ny = n;
        const float e = g.dx; g.dx = g.dy; g.dy = e;
        const float f = g.cx; g.cx = g.cy; g.cy = f;
    }

    // InitializeSiddon, cuda/siddon.cu:80-149
    const float dist = sqrtf((s.x - t.x) * (s.x - t.x) + (s.y - t.y) * (s.y - t.y) + (s.z - t.z) * (s.z - t.z));
    const float3 si = make_float3((s.x - g.cx) / g.dx + g.nx / 2.0f, (s.y - g.cy) / g.dy + g.ny / 2.0f,
                                  (s.z - g.cz) / g.dz + g.nz / 2.0f);
    const float3 ti = make_float3((t.x - g.cx) / g.dx + g.nx / 2.0f, (t.y - g.cy) / g.dy + g.ny / 2.0f,
                                  (t.z - g.cz) / g.dz + g.nz / 2.0f);
    const float dx = ti.x - si.x, dy = ti.y - si.y, dz = ti.z - si.z;
    const bool px = fabsf(dx) > kEps, py = fabsf(dy) > kEps, pz = fabsf(dz) > kEps;
    if ((!px && (si.x < 0 || si.x > g.nx)) || (!py && (si.y < 0 || si.y > g.ny)) ||
        (!pz && (si.z < 0 || si.z > g.nz)))
        return false;
    const float rdx = 1.0f / dx, rdy = 1.0f / dy, rdz = 1.0f / dz;
    float a0x = 1.f, aNx = 0.f, a0y = 1.f, aNy = 0.f, a0z = 1.f, aNz = 0.f;
    if (px) { a0x = -si.x * rdx; aNx = ((float)g.nx - si.x) * rdx; }
    if (py) { a0y = -si.y * rdy; aNy = ((float)g.ny - si.y) * rdy; }
    if (pz) { a0z = -si.z * rdz; aNz = ((float)g.nz - si.z) * rdz; }
    const float amin = fmaxf(0.f, fmaxf(fminf(a0x, aNx), fmaxf(fminf(a0y, aNy), fminf(a0z, aNz))));
    const float amax = fminf(1.f, fminf(fmaxf(a0x, aNx), fminf(fmaxf(a0y, aNy), fmaxf(a0z, aNz))));
    if (amax <= amin) return false;

    int ix = (int)(si.x + amin * dx), iy = (int)(si.y + amin * dy), iz = (int)(si.z + amin * dz);
    if (px) ix = min(max(ix, 0), g.nx - 1);
    if (py) iy = min(max(iy, 0), g.ny - 1);
    if (pz) iz = min(max(iz, 0), g.nz - 1);
    if (ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || iz < 0 || iz >= g.nz) return false;

    // first crossings and steps, scaled to mm
    r.ax = px ? ((float)(dx > kEps ? ix + 1 : ix) - si.x) * rdx * dist : 2.f * dist;
    r.ay = py ? ((float)(dy > kEps ? iy + 1 : iy) - si.y) * rdy * dist : 2.f * dist;
    r.az = pz ? ((float)(dz > kEps ? iz + 1 : iz) - si.z) * rdz * dist : 2.f * dist;
    r.dax = px ? fabsf(rdx) * dist : 0.f;
    r.day = py ? fabsf(rdy) * dist : 0.f;
    r.daz = pz ? fabsf(rdz) * dist : 0.f;
    const int nxy = g.nx * g.ny;
    // byte offsets into the volume (below 2 GiB, checked on the host)
    r.mx = dx > 0 ? 4 : -4;
    r.my = dy > 0 ? 4 * g.nx : -4 * g.nx;
    r.mz = dz > 0 ? 4 * nxy : -4 * nxy;
    // the traversal ends with the crossing that takes a counter below zero
    // (cuda/siddon.cu:169-228 stops when the index leaves the grid)
    r.rx = dx > 0 ? g.nx - 1 - ix : ix;
    r.ry = dy > 0 ? g.ny - 1 - iy : iy;
    r.rz = dz > 0 ? g.nz - 1 - iz : iz;
    r.a = amin * dist;
    r.off = 4 * ((iz * g.ny + iy) * g.nx + ix);
    return true;
}

// One step of the traversal: returns the segment length assigned to the voxel
// at r.off before the step and advances to the next voxel.
__device__ inline float siddon_step(SiddonRay& ray, bool& inside) {
    // The state is worked on as a local copy: selects between fields of a
    // referenced struct are otherwise folded into a load through a selected
    // address, which keeps the ray in memory (the compiler then places it in
    // LDS and every step waits on an LDS read).
    SiddonRay r = ray;
    const float anow = fminf(fminf(r.ax, r.ay), r.az);
    const bool bx = r.ax == anow;
    const bool by = !bx && r.ay == anow;
    const bool bz = !bx && !by;
    const float len = anow - r.a;
    r.a = anow;
    r.off += bx ? r.mx : (by ? r.my : r.mz);
    r.rx -= bx ? 1 : 0;
    r.ry -= by ? 1 : 0;
    r.rz -= bz ? 1 : 0;
    r.ax = bx ? r.ax + r.dax : r.ax;
    r.ay = by ? r.ay + r.day : r.ay;
    r.az = bz ? r.az + r.daz : r.az;
    inside = (r.rx | r.ry | r.rz) >= 0;
    ray = r;
    return len;
}

// Line integral of the volume behind 'vol' along a set-up ray.  Software
// pipelined: the voxel sequence does not depend on the voxel values, so NP
// steps are traced first and their NP gathers are issued together; steps past
// the exit read byte offset 0 with weight 0.
__device__ inline float siddon_integral(__amdgpu_buffer_rsrc_t vol, SiddonRay& r) {
    float val = 0.f;
    bool inside = true;
    constexpr int NP = 4;
    while (inside) {
        int id[NP];
        float len[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) {
            id[k] = inside ? r.off : 0;
            bool in_next;
            const float l = siddon_step(r, in_next);
            len[k] = inside ? l : 0.f;
            inside = inside && in_next;
        }
        float f[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) f[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(vol, id[k], 0, 0));
#pragma unroll
        for (int k = 0; k < NP; k++) val = fmaf(f[k], len[k], val);
    }
    return val;
}

// Detector cell centre of ray (iu, iv) of a view, cuda/siddonCone.cu:27-31.
__device__ inline float3 cell_centre(const float* pdc, const float* pdu, const float* pdv, int view, int iu, int iv,
                                     const Det& d) {
    const float u = ((float)iu - d.off_u - (float)(d.nu - 1) / 2.0f) * d.du;
    const float v = ((float)iv - d.off_v - (float)(d.nv - 1) / 2.0f) * d.dv;
    const float3 c = ld3(pdc, view), U = ld3(pdu, view), V = ld3(pdv, view);
    return make_float3(c.x + U.x * u + V.x * v, c.y + U.y * u + V.y * v, c.z + U.z * u + V.z * v);
}


}  // namespace ctp
