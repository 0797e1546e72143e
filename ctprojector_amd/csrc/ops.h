// Internal stream-ordered operations of the projector library.  All array
// arguments are device pointers unless named host; every function enqueues its
// work on 'stream' and may synchronise that stream only where noted.
#pragma once

#include "common.h"

#include <memory>
#include <vector>

namespace ctp {

struct ConeGeomDev {               // [nview][3] float rows (x, y, z), device
    const float* det_center;
    const float* det_u;
    const float* det_v;
    const float* src;
};
struct ConeGeomHost {              // the same rows in host memory
    const float* det_center;
    const float* det_u;
    const float* det_v;
    const float* src;
};

// Detector rows feeding an image slice: (1 - w1) * row r0 + w1 * row r1, or
// nothing when r0 < 0.
struct SliceRow {
    int r0, r1;
    float w1;
    int pad;
};

// sino[view][u][c] = rows(z) of prj[b][view][.][u] for c = b*nz + z < C, zero for
// C <= c < cs (channel-last layout shared by the pixel-driven backprojectors)
void slices_to_channel_last(float* sino, const float* prj, const SliceRow* rows_dev, int C, int cs, int nz,
                            const Det& d, int nview, hipStream_t stream);

// ---- Siddon cone beam (cuda/siddonCone.cu) ----
// prj += A img   (the reference's FP accumulates, cuda/siddon.cu:230)
void siddon_cone_fp(float* prj, const float* img, const ConeGeomDev& geo, int B, const Grid& g, const Det& d,
                    int nview, hipStream_t stream);
// img += A^T prj  (exact transpose; synchronises 'stream' at the end)
void siddon_cone_bp(float* img, const float* prj, const ConeGeomDev& geo, const ConeGeomHost& hgeo, int B,
                    const Grid& g, const Det& d, int nview, hipStream_t stream);

// ---- distance driven fan / parallel (cuda/distanceDriven2D*.cu) ----
// 'deg_host' are the view angles in host memory (per-view decisions are taken
// in FP64 on the host); prj is overwritten, img accumulates.
void dd2d_fp(float* prj, const float* img, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
             float dsd, float dso, int type, bool fan, hipStream_t stream);
void dd2d_bp(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
             float dsd, float dso, int type, bool fan, hipStream_t stream);
// dd2d_fp / dd2d_bp keep the geometry tables of recent (geometry, angles)
// pairs on the device; this frees the entries no call is using.
void release_dd_call_cache();
// dd2d_fp with its scratch (channel-last image copies and tables) in a
// file-scope device array of kDDFPNoAllocScratchBytes: allocates no device
// memory; fails with an error when the call needs more scratch; calls that
// overlap on one GPU race on the array.
constexpr size_t kDDFPNoAllocScratchBytes = size_t(80) << 20;
void dd2d_fp_noalloc(float* prj, const float* img, const float* deg_host, int B, const Grid& g, const Det& d,
                     int nview, float dsd, float dso, int type, bool fan, hipStream_t stream);
// img += A^T prj, the exact transpose of dd2d_fp (same type bits)
void dd2d_fp_adjoint(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d,
                     int nview, float dsd, float dso, int type, bool fan, hipStream_t stream);
// Persistent plan of one DD geometry and batch count: tables built once, calls
// launch kernels only on 'stream' (asynchronous, no host synchronisation).
// forward overwrites prj; backward and adjoint accumulate into img.
struct DDPlan;
DDPlan* dd_plan_create(const float* deg_host, int B, const Grid& g, const Det& d, int nview, float dsd, float dso,
                       int type, bool fan, hipStream_t stream);
void dd_plan_forward(DDPlan* p, float* prj, const float* img, hipStream_t stream);
void dd_plan_backward(DDPlan* p, float* img, const float* prj, hipStream_t stream);
void dd_plan_adjoint(DDPlan* p, float* img, const float* prj, hipStream_t stream);
void dd_plan_destroy(DDPlan* p);

// ---- batched plans of uneven DD jobs (no counterpart in the reference) ----
struct DDBatchJobSpec {
    bool fan;                 // equiangular fan beam (else parallel beam)
    int type;                 // typeProjector bits
    int B;                    // batch items
    Grid g;
    Det d;
    int nview;
    float dsd, dso;
    const float* deg;         // host [nview], read at creation
    float* img;               // device [B][nz][ny][nx], bound
    float* prj;               // device [B][nview][nv][nu], bound
};

class DDBatchPlan {
public:
    DDBatchPlan(const std::vector<DDBatchJobSpec>& jobs, bool autotune);
    ~DDBatchPlan();
    void forward(hipStream_t s);    // prj = A img for every job (overwrites)
    void backward(hipStream_t s);   // img = the reference's backprojection of prj, every job (overwrites)
    void tune();
    void config(int& fp_lanes, int& bp_lanes, int& fp_persistent, int& bp_persistent, int& fp_streams,
                int& bp_streams) const;
    int launches(bool forward_dir) const;   // kernel launches of one call (without the counter reset)
    struct Impl;
private:
    std::unique_ptr<Impl> impl_;
};

// ---- batched ramp filters of uneven jobs (no counterpart in the reference) ----
struct FilterBatchJobSpec {
    int B, nu, nv, nview;
    float da, off_u, dso;     // da: angular pitch (fan) or du (parallel)
    int filter;               // typeFilter of the single-call filter
    bool equispace;           // parallel beam
    const float* in;          // device [B][nview][nv][nu], bound
    float* out;               // device [B][nview][nv][nu], bound
};

class FilterBatchPlan {
public:
    explicit FilterBatchPlan(const std::vector<FilterBatchJobSpec>& jobs);
    ~FilterBatchPlan();
    void run(hipStream_t s);   // out = the single-call ramp filter of in, every job
    int launches() const;      // kernel launches of one call (one per transform size)
    struct Impl;
private:
    std::unique_ptr<Impl> impl_;
};

// ---- batched cone-beam plans (no counterpart in the reference) ----
struct ConeJobSpec {
    Grid g;
    Det d;
    int nview;
    const float *det_center, *det_u, *det_v, *src;   // host [nview][3], read at creation
    float* img;                                      // device [nz][ny][nx], bound
    float* prj;                                      // device [nview][nv][nu], bound
};

class ConePlan {
public:
    ConePlan(const std::vector<ConeJobSpec>& jobs, bool autotune);
    ~ConePlan();
    void forward(hipStream_t s);    // prj = A img for every job (overwrites)
    void backward(hipStream_t s);   // img = A^T prj for every job (overwrites)
    void tune();
    void config(int& tile_u, int& xcd_order, int& zrun) const;
    struct Impl;
private:
    std::unique_ptr<Impl> impl_;
};

// ---- detector rows -> slice channels ----
struct ChanRowHost {      // one detector row of a slice channel and its weight
    int iv;
    float w;
};
// q[view][u][c] = sum over the rows (iv, w) of channel c of w * prj[b][view][iv][u],
// c = b * nz + z (channels C..cs-1 zero); cstart/crow: CSR lists on the device
void rows_to_channel_last(float* q, const float* prj, const int* cstart, const ChanRowHost* crow, int C, int cs,
                          int nz, const Det& d, int nview, hipStream_t stream);

// ---- Siddon fan beam (cuda/siddonFan.cu); angles on the device ----
void siddon_fan_fp(float* prj, const float* img, const float* deg_dev, int B, const Grid& g, const Det& d, int nview,
                   float dsd, float dso, hipStream_t stream);
void siddon_fan_bp(float* img, const float* prj, const float* deg_dev, int B, const Grid& g, const Det& d, int nview,
                   float dsd, float dso, hipStream_t stream);

// ---- pixel-driven backprojection helpers shared by dd2d.hip and fbp.hip ----
// view groups for a launch over npix pixels x nchunk channel chunks
int bp_view_groups(size_t npix, int nchunk, int nview);
// img[i] += sum_k part[k * n + i], k = 0 .. groups-1 in order
void reduce_parts(float* img, const float* part, int groups, size_t n, hipStream_t stream);

// Stores a lane's CH channel accumulators: img += acc (one view group) or the
// group's partial image (several groups, summed by reduce_parts).
template <int CH>
__device__ inline void store_bp_result(float* __restrict__ img, float* __restrict__ part, const float* acc, int grp,
                                       int c0, int C, size_t npix, size_t pix) {
#pragma unroll
    for (int j = 0; j < CH; j++) {
        const int cc = c0 + j;
        if (cc >= C) continue;
        if (part) part[((size_t)grp * C + cc) * npix + pix] = acc[j];
        else img[(size_t)cc * npix + pix] += acc[j];
    }
}

// ---- distance driven tomosynthesis (cuda/distanceDriven3D*.cu) ----
// geometry rows on device; outputs are overwritten (the reference writes '=')
void tomo_fp(float* prj, const float* img, const float* det_center, const float* src, int B, const Grid& g,
             const Det& d, int nview, int type, hipStream_t stream);
void tomo_bp(float* img, const float* prj, const float* det_center, const float* src, int B, const Grid& g,
             const Det& d, int nview, int type, hipStream_t stream);

// ---- ramp filter and fan FBP backprojection (cuda/fbp*.cu) ----
// out is overwritten with the filtered projections
void ramp_filter(float* out, const float* prj, int B, int nu, int nv, int nview, float da, float off_u, float dsd,
                 float dso, int filter, bool equispace, hipStream_t stream);
// tomosynthesis FDK filter (cuda/fbpTomo.cu); geometry rows on the host
void tomo_filter(float* out, const float* prj, const float* det_center_host, const float* src_host, int B, int nu,
                 int nv, int nview, float du, float dx, float dz, int filter, float cutx, float cutz,
                 hipStream_t stream);
// img += FBP-weighted pixel-driven fan backprojection (fan), or the parallel
// pixel-driven backprojection (fan = false; dsd, dso unused)
void fan_fbp_bp(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
                float dsd, float dso, hipStream_t stream, bool fan = true);

// ---- anisotropic TV prior for SQS (prior-cuda/TotalVariation.cu) ----
void tv_sqs(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz, int B, int nx, int ny,
            int nz, float eps, hipStream_t stream);
void tv_sqs_update(float* out, float* var, const float* img, const float* bp, const float* norm, float beta, float acc,
                   float wx, float wy, float wz, int B, int nx, int ny, int nz, float eps, hipStream_t stream);
// dimg = vector-Jacobian product of (s1, s2, var) of tv_sqs with (g1, g2, gv);
// gv may be null (zero); 8 B of scratch per voxel: 'scratch', or the stream's workspace when it is null
void tv_sqs_backward(float* dimg, const float* img, const float* g1, const float* g2, const float* gv, float wx,
                     float wy, float wz, int B, int nx, int ny, int nz, float eps, hipStream_t stream,
                     float* scratch = nullptr);

}  // namespace ctp
