/*
 * ctprojector.h - C ABI of libprojector.so (MI355X / gfx950 build).
 *
 * Drop-in replacement for the reference's libprojector.so of wudufan/CTProjector
 * (src/ct_projector/kernel/projector/cuda/, abbreviated cuda/ below).  Every
 * function keeps the reference's name, argument order and meaning; the
 * reference file:line of the function it replaces is given above each
 * declaration.
 *
 * Conventions (as in the reference):
 *   images       float img[B][nz][ny][nx]
 *   projections  float prj[B][nview][nv][nu]
 *   geometry     float [nview][3] rows (x, y, z); x is the fastest image axis
 *   angles       float [nview], radians
 *   fan-beam calls take the angular pitch da = du / dsd in place of du
 *   c*     functions take host pointers and stage through device memory;
 *   cupy*  functions take device pointers (any HIP device memory, e.g. torch
 *          tensors); where the reference accumulates (Siddon FP, every BP), the
 *          caller passes a zero-initialised output.
 * All calls run on the current HIP device, return when the results are
 * complete, and return 0 on success or a non-zero hipError_t code.  A failure
 * prints "<function>() failed: <reason>" to stderr (as the reference does) but
 * never aborts the process; ctpGetLastErrorMessage() returns that text.
 *
 * typeProjector of the fan/parallel distance-driven projector is the
 * reference's bit mask (cuda/distanceDriven2D.cu:12-25): bit0 FBP weighting,
 * bit1 branchless, bit2 force FBP weighting.  For the tomosynthesis projector
 * typeProjector 1 selects the branchless mode (cuda/distanceDriven3D.cu:27,77).
 * Filter codes: 0 Ram-Lak, 1 Hamming, 2 Hann, 3 cosine (cuda/fbp.h:5-7).
 */
#ifndef CTPROJECTOR_H
#define CTPROJECTOR_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cuda/projector.cu:12 */
int SetDevice(int device);

/* ---- Siddon cone beam, arbitrary trajectory ------------------------------ */
/* cuda/siddonCone.cu:137  prj = A img (host arrays) */
int cSiddonConeProjectionArbitrary(float* prj, const float* img, const float* detCenter, const float* detU,
                                   const float* detV, const float* src, size_t nBatches, size_t nx, size_t ny,
                                   size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                                   size_t nv, size_t nview, float du, float dv, float off_u, float off_v);
/* cuda/siddonCone.cu:233  prj += A img (device arrays) */
int cupySiddonConeProjectionArbitrary(float* prj, const float* img, const float* detCenter, const float* detU,
                                      const float* detV, const float* src, size_t nBatches, size_t nx, size_t ny,
                                      size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                                      size_t nv, size_t nview, float du, float dv, float off_u, float off_v);
/* cuda/siddonCone.cu:282  img = A^T prj (host arrays; exact transpose of the FP) */
int cSiddonConeBackprojectionArbitrary(float* img, const float* prj, const float* detCenter, const float* detU,
                                       const float* detV, const float* src, size_t nBatches, size_t nx, size_t ny,
                                       size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                                       size_t nu, size_t nv, size_t nview, float du, float dv, float off_u,
                                       float off_v);
/* cuda/siddonCone.cu:378  img += A^T prj (device arrays) */
int cupySiddonConeBackprojectionArbitrary(float* img, const float* prj, const float* detCenter, const float* detU,
                                          const float* detV, const float* src, size_t nBatches, size_t nx, size_t ny,
                                          size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                                          size_t nu, size_t nv, size_t nview, float du, float dv, float off_u,
                                          float off_v);

/* ---- distance driven, fan beam (equiangular) and parallel beam ----------- */
/* cuda/distanceDriven2D.cu:109 */
int cupyDistanceDrivenFanProjection(float* prj, const float* img, const float* deg, size_t nBatches, size_t nx,
                                    size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                                    size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v,
                                    float dsd, float dso, int typeProjector);
/* No counterpart in the reference.  cupyDistanceDrivenFanProjection for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same arguments and
 * results, with the call's scratch (two channel-last copies of the B*nz slices
 * plus view/row tables) in a fixed 80 MiB device array per GPU, held until the
 * process ends.  A call needing more scratch returns an error.  Calls that
 * overlap on one GPU (other threads or streams) race on the array. */
int ctpDistanceDrivenFanProjectionNoAlloc(float* prj, const float* img, const float* deg, size_t nBatches, size_t nx,
                                          size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                                          float cz, size_t nu, size_t nv, size_t nview, float da, float dv,
                                          float off_a, float off_v, float dsd, float dso, int typeProjector);
/* cuda/distanceDriven2D.cu:158 */
int cupyDistanceDrivenParallelProjection(float* prj, const float* img, const float* deg, size_t nBatches, size_t nx,
                                         size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                                         float cz, size_t nu, size_t nv, size_t nview, float da, float dv,
                                         float off_a, float off_v, int typeProjector);
/* cuda/distanceDriven2D.cu:205 */
int cDistanceDrivenFanProjection(float* prj, const float* img, const float* deg, size_t nBatches, size_t nx,
                                 size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                                 size_t nu, size_t nv, size_t nview, float da, float dv, float off_a, float off_v,
                                 float dsd, float dso, int typeProjector);
/* cuda/distanceDriven2D.cu:285 */
int cDistanceDrivenParallelProjection(float* prj, const float* img, const float* deg, size_t nBatches, size_t nx,
                                      size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                                      float cz, size_t nu, size_t nv, size_t nview, float da, float dv, float off_a,
                                      float off_v, int typeProjector);
/* cuda/distanceDriven2D.cu:384 */
int cupyDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg, size_t nBatches, size_t nx,
                                        size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                                        float cz, size_t nu, size_t nv, size_t nview, float da, float dv,
                                        float off_a, float off_v, float dsd, float dso, int typeProjector);
/* cuda/distanceDriven2D.cu:434 */
int cupyDistanceDrivenParallelBackprojection(float* img, const float* prj, const float* deg, size_t nBatches,
                                             size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx,
                                             float cy, float cz, size_t nu, size_t nv, size_t nview, float da,
                                             float dv, float off_a, float off_v, int typeProjector);
/* cuda/distanceDriven2D.cu:482 */
int cDistanceDrivenFanBackprojection(float* img, const float* prj, const float* deg, size_t nBatches, size_t nx,
                                     size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                                     size_t nu, size_t nv, size_t nview, float da, float dv, float off_a,
                                     float off_v, float dsd, float dso, int typeProjector);
/* cuda/distanceDriven2D.cu:560 */
int cDistanceDrivenParallelBackprojection(float* img, const float* prj, const float* deg, size_t nBatches,
                                          size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx,
                                          float cy, float cz, size_t nu, size_t nv, size_t nview, float da, float dv,
                                          float off_a, float off_v, int typeProjector);

/* No counterpart in the reference (whose backprojector is not the transpose of
 * its forward projector): img += A^T prj for the fan / parallel forward
 * projector above with the same typeProjector bits, device arrays; the exact
 * gradient of the forward projection. */
int ctpDistanceDrivenFanProjectionAdjoint(float* img, const float* prj, const float* deg, size_t nBatches, size_t nx,
                                          size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy,
                                          float cz, size_t nu, size_t nv, size_t nview, float da, float dv, float off_a,
                                          float off_v, float dsd, float dso, int typeProjector);
int ctpDistanceDrivenParallelProjectionAdjoint(float* img, const float* prj, const float* deg, size_t nBatches,
                                               size_t nx, size_t ny, size_t nz, float dx, float dy, float dz, float cx,
                                               float cy, float cz, size_t nu, size_t nv, size_t nview, float da,
                                               float dv, float off_a, float off_v, int typeProjector);

/* No counterpart in the reference (whose every call rebuilds its tables and
 * synchronises, cuda/distanceDriven2D.cu:109-634): a persistent plan of one
 * fan (fan = 1) or parallel (fan = 0; dsd, dso unused) geometry and batch
 * count.  deg: host array of nview angles; the other arguments as in
 * cupyDistanceDrivenFanProjection.  Forward writes prj = A img (as
 * cupyDistanceDrivenFanProjection), Backward adds the reference's
 * backprojection (as cupyDistanceDrivenFanBackprojection), Adjoint adds the
 * exact transpose (as ctpDistanceDrivenFanProjectionAdjoint); device arrays,
 * asynchronous on 'stream' (a hipStream_t; NULL = the null stream), no host
 * synchronisation.  Each direction keeps its scratch in the plan, so the
 * calls queue kernels only and can be captured in a HIP graph; calls of one
 * direction on one plan must not overlap in time (issue them on one stream,
 * or order them with events).  The plan must be used on the device current
 * at creation; Destroy waits for that device. */
int ctpDDPlanCreate(void** plan, int fan, const float* deg, size_t nBatches, size_t nx, size_t ny, size_t nz,
                    float dx, float dy, float dz, float cx, float cy, float cz, size_t nu, size_t nv, size_t nview,
                    float da, float dv, float off_a, float off_v, float dsd, float dso, int typeProjector);
int ctpDDPlanForward(void* plan, float* prj, const float* img, void* stream);
int ctpDDPlanBackward(void* plan, float* img, const float* prj, void* stream);
int ctpDDPlanAdjoint(void* plan, float* img, const float* prj, void* stream);
int ctpDDPlanDestroy(void* plan);

/* ---- Siddon fan beam (equiangular arc, stacked slices; host arrays) ------- */
/* cuda/siddonFan.cu:128  prj += A img; rows are horizontal planes z = (iv - (nv-1)/2 + off_v) dv */
int cSiddonFanProjection(float* prj, const float* img, const float* deg, int nBatches, size_t nx, size_t ny,
                         size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu, size_t nv,
                         size_t nview, float da, float dv, float off_a, float off_v, float dsd, float dso);
/* cuda/siddonFan.cu:202  img += A^T prj (exact transpose, deterministic; the reference uses atomics) */
int cSiddonFanBackprojection(float* img, const float* prj, const float* deg, int nBatches, size_t nx, size_t ny,
                             size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                             size_t nv, size_t nview, float da, float dv, float off_a, float off_v, float dsd,
                             float dso);

/* ---- distance driven tomosynthesis --------------------------------------- */
/* cuda/distanceDriven3D.cu:184 */
int cDistanceDrivenTomoProjection(float* prj, const float* img, const float* detCenter, const float* src,
                                  size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                  float cx, float cy, float cz, size_t nu, size_t nv, size_t nview, float du,
                                  float dv, float off_u, float off_v, int typeProjector);
/* cuda/distanceDriven3D.cu:264 */
int cupyDistanceDrivenTomoProjection(float* prj, const float* img, const float* detCenter, const float* src,
                                     size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                     float cx, float cy, float cz, size_t nu, size_t nv, size_t nview, float du,
                                     float dv, float off_u, float off_v, int typeProjector);
/* cuda/distanceDriven3D.cu:312 */
int cDistanceDrivenTomoBackprojection(float* img, const float* prj, const float* detCenter, const float* src,
                                      size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                                      float cx, float cy, float cz, size_t nu, size_t nv, size_t nview, float du,
                                      float dv, float off_u, float off_v, int typeProjector);
/* cuda/distanceDriven3D.cu:393 */
int cupyDistanceDrivenTomoBackprojection(float* img, const float* prj, const float* detCenter, const float* src,
                                         size_t nBatches, size_t nx, size_t ny, size_t nz, float dx, float dy,
                                         float dz, float cx, float cy, float cz, size_t nu, size_t nv, size_t nview,
                                         float du, float dv, float off_u, float off_v, int typeProjector);

/* ---- ramp filters and fan FBP backprojection ----------------------------- */
/* cuda/fbp.cu:165 */
int cupyfbpFanFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview, float da,
                     float dv, float off_a, float off_v, float dsd, float dso, int typeFilter);
/* cuda/fbp.cu:204 */
int cfbpFanFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview, float da,
                  float dv, float off_a, float off_v, float dsd, float dso, int typeFilter);
/* cuda/fbpParallel.cu:111 */
int cupyfbpParallelFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview, float du,
                          float dv, float off_u, float off_v, int typeFilter);
/* cuda/fbpParallel.cu:146 */
int cfbpParallelFilter(float* pFPrj, const float* pPrj, int nBatches, size_t nu, size_t nv, size_t nview, float du,
                       float dv, float off_u, float off_v, int typeFilter);
/* cuda/fbp.cu:362 */
int cupyfbpFanBackprojection(float* pImg, const float* pPrj, const float* pDeg, size_t nBatches, size_t nx, size_t ny,
                             size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu,
                             size_t nv, size_t nview, float da, float dv, float off_a, float off_v, float dsd,
                             float dso);
/* cuda/fbp.cu:409 */
int cfbpFanBackprojection(float* pImg, const float* pPrj, const float* pDeg, size_t nBatches, size_t nx, size_t ny,
                          size_t nz, float dx, float dy, float dz, float cx, float cy, float cz, size_t nu, size_t nv,
                          size_t nview, float da, float dv, float off_a, float off_v, float dsd, float dso);

/* ---- TV prior for SQS (the reference's libprior.so) ---------------------- */
/* prior-cuda/TotalVariation.cu:182  s1, s2, var of the anisotropic TV surrogate
 * (device arrays [B][nz][ny][nx]); wx, wy, wz weight the x, y, z differences */
int cupyTVSQS3D(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz, size_t nBatches,
                size_t nx, size_t ny, size_t nz, float eps);
/* prior-cuda/TotalVariation.cu:217  the same on host arrays (every batch item
 * processed; the reference allocates one batch item's worth, :231-246) */
int cTVSQS3D(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz, size_t nBatches,
             size_t nx, size_t ny, size_t nz, float eps);
/* No counterpart in the reference (it composes tv_sqs and array arithmetic in
 * Python, recon/cupy/sqs_algorithms.py:40-48): one fused SQS step,
 * out = img - (acc*bp + beta*s1(img)) / (norm + beta*s2(img)); var(img) written
 * when var is not NULL.  out must not alias img.  Device arrays. */
int ctpTVSQSUpdate3D(float* out, float* var, const float* img, const float* bp, const float* norm, float beta,
                     float acc, float wx, float wy, float wz, size_t nBatches, size_t nx, size_t ny, size_t nz,
                     float eps);
/* No counterpart in the reference (gradients for unrolled, differentiable
 * SQS-TV networks): dimg = J_s1^T g1 + J_s2^T g2 + J_var^T gvar, the
 * vector-Jacobian product of cupyTVSQS3D's outputs at img.  gvar may be NULL
 * (zero).  dimg is overwritten and must not alias the inputs.  Device arrays. */
int ctpTVSQSBackward3D(float* dimg, const float* img, const float* g1, const float* g2, const float* gvar, float wx,
                       float wy, float wz, size_t nBatches, size_t nx, size_t ny, size_t nz, float eps);
/* No counterpart in the reference: cupyTVSQS3D and ctpTVSQSBackward3D queued
 * on 'stream' (a hipStream_t, NULL for the null stream) without waiting for
 * completion, for callers that order their work by stream (the
 * differentiable SQS-TV network).  Arguments and results as above;
 * ctpTVSQSBackward3DAsync uses 'scratch' (2 floats per voxel, 8-byte
 * aligned, distinct from the other arrays) or, when it is NULL, the
 * library's workspace of that stream and host thread; with caller scratch the
 * call queues kernels only and can be captured in a HIP graph. */
int ctpTVSQS3DAsync(float* s1, float* s2, float* var, const float* img, float wx, float wy, float wz, size_t nBatches,
                    size_t nx, size_t ny, size_t nz, float eps, void* stream);
int ctpTVSQSBackward3DAsync(float* dimg, const float* img, const float* g1, const float* g2, const float* gvar,
                            float* scratch, float wx, float wy, float wz, size_t nBatches, size_t nx, size_t ny,
                            size_t nz, float eps, void* stream);

/* ---- batched cone-beam plans (no counterpart in the reference) ------------ */
/* Many Siddon cone-beam jobs (volumes of different sizes, each with its own
 * trajectory and detector) computed by one persistent launch per direction.
 * ctpConePlanCreate reads the host geometry, prepares every geometry-only
 * table on the device and binds the jobs' device buffers img and prj (which
 * must stay allocated while the plan lives).  Forward (prj = A img) and
 * Backward (img = A^T prj) overwrite their outputs and only launch kernels on
 * 'stream' (a hipStream_t, NULL for the null stream): no allocation, copy,
 * memset or host synchronisation, so they can be captured into a HIP graph.
 * They return before the work completes.  Jobs need the column geometry
 * (det_v along z, det_u horizontal in every view).  autotune != 0 times the
 * launch shapes on the bound buffers (overwriting them) and keeps the fastest.
 * ctpConePlanConfig reports them: forward tile width (0 = column-driven form),
 * block order (1 = XCD bands), backprojection z-run.  Calls on one plan must
 * not overlap in time (they share its work counters and tables): issue them on
 * one stream, or order them with events. */
typedef struct ctpConeJob {
    size_t nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
    size_t nu, nv, nview;
    float du, dv, off_u, off_v;
    const float* det_center; /* host [nview][3] */
    const float* det_u;
    const float* det_v;
    const float* src;
    float* img;              /* device [nz][ny][nx] */
    float* prj;              /* device [nview][nv][nu] */
} ctpConeJob;
int ctpConePlanCreate(void** plan, const ctpConeJob* jobs, int njobs, int autotune);
int ctpConePlanForward(void* plan, void* stream);
int ctpConePlanBackward(void* plan, void* stream);
int ctpConePlanConfig(void* plan, int* tile_u, int* xcd_order, int* zrun);
int ctpConePlanDestroy(void* plan);

/* No counterpart in the reference (whose DD entry points take one geometry and
 * one size per call, cuda/distanceDriven2D.cu:109-634): a batch of fan-beam
 * (fan = 1) or parallel-beam (fan = 0) distance-driven jobs of different
 * geometry, size, view count and batch count, bound to their device arrays.
 * Per job the arguments mean what they mean for
 * cupyDistanceDrivenFanProjection (da: angular pitch for the fan beam, du for
 * the parallel beam; dsd, dso unused for the parallel beam); deg is a host
 * array read at creation.  Forward writes prj = A img for every job, Backward
 * writes img = the reference's backprojection of prj (the job's
 * typeProjector bits) for every job; each job's result is the single call's.
 * The jobs whose kernels are the same template instance run in one launch,
 * and the launches of different kernel classes may run concurrently on up to
 * four internal streams forked from and joined to 'stream' (Config reports
 * the launch shapes, stream counts and launch counts; lanes 0: each job's
 * own default, as in the single calls).  Autotune times the launch shapes on
 * the bound arrays at creation (prj and img are overwritten).  Calls are
 * asynchronous on 'stream' and launch kernels only, so they can be captured
 * in a HIP graph; the batch must be used on the device current at creation.
 * Calls on one batch must not overlap in time (they share its scratch, work
 * counters and internal streams): issue them on one stream, or order them
 * with events. */
typedef struct ctpDDJob {
    int fan;
    int typeProjector;
    size_t nBatches, nx, ny, nz;
    float dx, dy, dz, cx, cy, cz;
    size_t nu, nv, nview;
    float da, dv, off_a, off_v;
    float dsd, dso;
    const float* deg;        /* host [nview] */
    float* img;              /* device [nBatches][nz][ny][nx] */
    float* prj;              /* device [nBatches][nview][nv][nu] */
} ctpDDJob;
int ctpDDBatchCreate(void** batch, const ctpDDJob* jobs, int njobs, int autotune);
int ctpDDBatchForward(void* batch, void* stream);
int ctpDDBatchBackward(void* batch, void* stream);
int ctpDDBatchConfig(void* batch, int* fp_lanes, int* bp_lanes, int* fp_persistent, int* bp_persistent,
                     int* fp_streams, int* bp_streams, int* fp_launches, int* bp_launches);
int ctpDDBatchDestroy(void* batch);

/* No counterpart in the reference (whose filters take one geometry per call,
 * cuda/fbp.cu:97-104, cuda/fbpParallel.cu:59): the ramp filters of a batch of
 * fan-beam (fan = 1: da the angular pitch, cupyfbpFanFilter) or parallel-beam
 * (fan = 0: da = du, dsd and dso unused, cupyfbpParallelFilter) jobs of
 * different detectors, view counts and batch counts, bound to their device
 * arrays (prj in, fprj out, [nBatches][nview][nv][nu]).  Run writes every
 * job's filtered sinogram, equal to its single call's, in one launch per FFT
 * size (nu up to 2048) plus one per job of 2049 to 5000 columns (the single
 * call's direct convolution), and only launches
 * kernels on 'stream' (capturable into a HIP graph; returns before the work
 * completes).  With ctpDDBatch* jobs of typeProjector bit 0 (FBP weights)
 * reading fprj, uneven FBP reconstructions take a few launches in all.
 * Calls on one batch must not overlap in time. */
typedef struct ctpFilterJob {
    int fan;
    int typeFilter;
    size_t nBatches, nu, nv, nview;
    float da, off_a, dsd, dso;
    const float* prj;        /* device */
    float* fprj;             /* device */
} ctpFilterJob;
int ctpFilterBatchCreate(void** batch, const ctpFilterJob* jobs, int njobs);
int ctpFilterBatchRun(void* batch, void* stream);
int ctpFilterBatchDestroy(void* batch);

/* cuda/fbpParallel.cu:273  parallel-beam pixel-driven backprojection (host arrays):
 * img += sum over views of the bilinear (truncating) sample of prj at
 * u = -((x cos + y sin)/du + off_u) + (nu-1)/2; typeProjector is unused */
int cfbpParallelBackprojection(float* pImg, const float* pPrj, const float* pDeg, size_t nBatches, size_t nx,
                               size_t ny, size_t nz, float dx, float dy, float dz, float cx, float cy, float cz,
                               size_t nu, size_t nv, size_t nview, float du, float dv, float off_u, float off_v,
                               int typeProjector);

/* cuda/fbpTomo.cu:386  tomosynthesis FDK filter (device arrays): per view a ramp
 * for the tilted virtual pitch du cos(b) with a Hamming (1) / Hann (2) window
 * of the x and z frequencies (cutoffs in (0, 1]), other codes Ram-Lak */
int cupyFbpTomoFilter(float* fprj, const float* prj, const float* detCenter, const float* src, int nBatches,
                      size_t nu, size_t nv, size_t nview, float du, float dx, float dz, int typeFilter, float cutoffX,
                      float cutoffZ);

/* ---- library services (no counterpart in the reference) ------------------ */
/* Text of the calling thread's last failure ("" if none). */
const char* ctpGetLastErrorMessage(void);
/* Sets the HIP stream (hipStream_t, NULL = null stream) on which the calling
 * thread's subsequent calls run; each call then synchronises only that stream,
 * so threads with their own streams overlap on the device. */
int ctpSetStream(void* stream);
/* Frees the device scratch memory the library keeps per (device, stream) and
 * the cached geometry tables of the distance-driven entry points. */
int ctpReleaseWorkspace(void);
/* Launch-shape knobs used by the autotuner ("cone_fp_tile_u", "cone_bp_zrun");
 * 0 restores the built-in default; unknown keys return hipErrorInvalidValue. */
int ctpSetTuning(const char* key, int value);
int ctpGetTuning(const char* key);
/* Per-kernel timing with HIP events on the launching stream (measurement). */
int ctpSetKernelTiming(int enable);
int ctpResetKernelTimes(void);
int ctpGetKernelTime(const char* name, double* total_ms, long* count);
int ctpKernelTimeReport(char* buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* CTPROJECTOR_H */
