// Ramp filtering (fan: equiangular, parallel: equispace) and the pixel-driven
// fan FBP backprojection on MI355X (gfx950): cuda/fbpMath.cu, cuda/fbp.cu,
// cuda/fbpParallel.cu of the reference.
//
// Ramp filter.  The reference zero-pads each detector row to L = 2nu-1, applies
// the windowed ramp in the frequency domain with cuFFT and reads the result at
// offset nu-1 (cuda/fbp.cu:106-134).  Because the windowed kernel h_w is
// defined on exactly L = 2nu-1 samples, that is the exact linear convolution
//     y[j] = scale * sum_m x[m] w[m] h_w[nu-1 + j - m]
// (no wrap-around), which is what is computed here, directly, in FP32 with the
// kernel built once in FP64 on the host.  h_w for the Hamming and Hann windows
// is a 3-tap circular smoothing of the spatial Ram-Lak kernel; the cosine
// window is applied with an explicit DFT in FP64.
//
// Fan FBP backprojection (bpFanKernel3D, cuda/fbp.cu:264-359): pixel driven;
// the bilinear detector interpolation is split into a per-slice row
// interpolation done once per view in a prep pass (channel-last layout) and a
// per-pixel column interpolation shared by all slices.
#include "common.h"
#include "runtime.h"
#include "ops.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace ctp {

// ---------------------------------------------------------------------------
// windowed ramp kernel (GetRamp, cuda/fbpMath.cu:90-268), FP64
// ---------------------------------------------------------------------------
static std::vector<double> ramp_kernel_fp64(int nu, double da, int filter, bool equispace) {
    const int L = 2 * nu - 1;
    const double PI = (double)kPI;
    std::vector<double> r(L), h(L);
    for (int i = 0; i < L; i++) {
        const int k = i - (nu - 1);
        if (k == 0) r[i] = 1.0 / (4 * da * da);
        else if (k % 2 != 0) {
            if (equispace) r[i] = -1.0 / (PI * PI * (double)k * k * da * da);
            else { const double sk = std::sin(k * da); r[i] = -1.0 / (PI * PI * sk * sk); }
        } else r[i] = 0;
    }
    if (filter == FILTER_HAMMING || filter == FILTER_HANN) {
        // W[f] = A + B cos(2 pi f / L) multiplies the DFT  <=>  circular taps {A, B/2, B/2}
        const double A = filter == FILTER_HAMMING ? (double)0.54f : (double)0.5f;
        const double Bw = filter == FILTER_HAMMING ? (double)0.46f : (double)0.5f;
        for (int i = 0; i < L; i++) h[i] = A * r[i] + 0.5 * Bw * (r[(i - 1 + L) % L] + r[(i + 1) % L]);
    } else if (filter == FILTER_COSINE) {
        std::vector<double> cs(L), sn(L);
        for (int k = 0; k < L; k++) { cs[k] = std::cos(2 * M_PI * k / L); sn[k] = std::sin(2 * M_PI * k / L); }
        std::vector<std::complex<double>> F(L);
        for (int f = 0; f < L; f++) {
            double re = 0, im = 0;
            for (int i = 0; i < L; i++) {
                const int k = (int)(((long)f * i) % L);
                re += r[i] * cs[k];
                im -= r[i] * sn[k];
            }
            const double w = std::fabs(std::cos(PI * f / (double)L));
            F[f] = std::complex<double>(re * w, im * w);
        }
        for (int i = 0; i < L; i++) {
            double re = 0;
            for (int f = 0; f < L; f++) {
                const int k = (int)(((long)f * i) % L);
                re += F[f].real() * cs[k] - F[f].imag() * sn[k];
            }
            h[i] = re / L;
        }
    } else {
        h = r;
    }
    return h;
}

static std::vector<float> cached_ramp(int nu, float da, int filter, bool equispace) {
    static std::mutex mu;
    static std::map<std::tuple<int, float, int, bool>, std::vector<float>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(nu, da, filter, equispace);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double> h = ramp_kernel_fp64(nu, (double)da, filter, equispace);
    std::vector<float> hf(h.begin(), h.end());
    if (cache.size() > 64) cache.clear();
    cache[key] = hf;
    return hf;
}

// ---------------------------------------------------------------------------
// direct linear convolution: RB rows per block, 4 outputs per thread
// ---------------------------------------------------------------------------
// nkern > 1: one kernel (and scale) per view, kernel of row r = (r / rows_per_kernel)
// mod nkern (tomosynthesis filter; RB = 1 so that a block never spans two views).
template <int RB>
__global__ __launch_bounds__(256) void ramp_conv_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                        const float* __restrict__ h, const float* __restrict__ pw,
                                                        long nrows, int nu, float scale,
                                                        const float* __restrict__ kscale, int rows_per_kernel,
                                                        int nkern) {
    extern __shared__ float sm[];
    const int L = 2 * nu - 1;
    float* hs = sm;                 // [L + 4]
    float* xs = sm + L + 4;         // [RB][nu]
    const long r0 = (long)blockIdx.x * RB;
    const int kk = nkern > 1 ? (int)((r0 / rows_per_kernel) % nkern) : 0;
    if (kscale) scale = kscale[kk];
    h += (size_t)kk * L;
    for (int i = threadIdx.x; i < L + 4; i += blockDim.x) hs[i] = i < L ? h[i] : 0.f;
    for (int rr = 0; rr < RB; rr++) {
        const long r = r0 + rr;
        for (int i = threadIdx.x; i < nu; i += blockDim.x)
            xs[rr * nu + i] = (r < nrows) ? in[r * nu + i] * pw[i] : 0.f;
    }
    __syncthreads();
    for (int j0 = threadIdx.x * 4; j0 < nu; j0 += blockDim.x * 4) {
        float acc[RB][4];
#pragma unroll
        for (int rr = 0; rr < RB; rr++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[rr][q] = 0.f;
        // h index for output j0+q and input m: nu-1 + j0 + q - m
        int base = nu - 1 + j0;   // m = 0
        float h0 = hs[base], h1 = hs[base + 1], h2 = hs[base + 2], h3 = hs[base + 3];
        for (int m = 0; m < nu; m++) {
#pragma unroll
            for (int rr = 0; rr < RB; rr++) {
                const float x = xs[rr * nu + m];
                acc[rr][0] = fmaf(x, h0, acc[rr][0]);
                acc[rr][1] = fmaf(x, h1, acc[rr][1]);
                acc[rr][2] = fmaf(x, h2, acc[rr][2]);
                acc[rr][3] = fmaf(x, h3, acc[rr][3]);
            }
            // slide the window: next m needs hs[base-1 .. base+2]
            h3 = h2; h2 = h1; h1 = h0;
            base--;
            h0 = base >= 0 ? hs[base] : 0.f;
        }
#pragma unroll
        for (int rr = 0; rr < RB; rr++) {
            const long r = r0 + rr;
            if (r >= nrows) break;
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (j0 + q < nu) out[r * nu + j0 + q] = acc[rr][q] * scale;
        }
    }
}

// ---------------------------------------------------------------------------
// FFT convolution: the same linear convolution as ramp_conv_kernel through a
// circular one of length N = 2^p >= 2nu - 1 (no wrap-around reaches the
// outputs 0..nu-1).  With g[k mod N] = h_w[nu - 1 + k] for |k| < nu, zero
// elsewhere, y = IFFT(FFT(x w) * G) on [0, nu), G = FFT(g).  h_w is even
// (the Ram-Lak taps are, and both window forms keep it), so G is real; it is
// built in FP64 on the host with scale / N folded in.  Two detector rows a, b
// travel as one complex row a + ib: a real G keeps them apart, and the
// inverse transform is taken as conj(FFT(conj(Z) G)), so one forward FFT
// serves both directions.  One 256-lane block per row pair; the row pair
// lives in LDS (N complex values) through mixed-radix Stockham passes
// (radix 8, then one radix-4 or radix-2 pass), twiddles from an FP64-built
// table.  Work per sample ~2 x 5 log2 N / 2 flops instead of 2 nu.
// ---------------------------------------------------------------------------
__device__ inline float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 caddf(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csubf(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (-i)
__device__ inline float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// forward DFT of R points in place, natural order: X[k] = sum_r v[r] e^{-2 pi i r k / R}
template <int R>
__device__ inline void dft_small(float2* v) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = caddf(a, b);
        v[1] = csubf(a, b);
    } else if constexpr (R == 4) {
        const float2 t0 = caddf(v[0], v[2]), t1 = csubf(v[0], v[2]);
        const float2 t2 = caddf(v[1], v[3]), t3 = cmul_mi(csubf(v[1], v[3]));
        v[0] = caddf(t0, t2);
        v[2] = csubf(t0, t2);
        v[1] = caddf(t1, t3);
        v[3] = csubf(t1, t3);
    } else {
        static_assert(R == 8, "radix 2, 4 or 8");
        float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4>(e);
        dft_small<4>(o);
        const float c = 0.70710678118654752f;
        o[1] = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));    // * e^{-i pi/4}
        o[2] = cmul_mi(o[2]);                                                // * e^{-i pi/2}
        o[3] = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));   // * e^{-3i pi/4}
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = caddf(e[k], o[k]);
            v[k + 4] = csubf(e[k], o[k]);
        }
    }
}

// One Stockham pass of radix R over buf[N] (Ns = product of the earlier radices):
// butterfly j reads buf[j + r N/R], twiddles by W_N^{r k N/(Ns R)} (k = j mod Ns),
// transforms, and writes buf[(j - k) R + k + r Ns].  At most 2 butterflies per
// lane (N / R <= 512).
template <int R>
__device__ inline void fft_pass(float2* buf, int N, int Ns, const float2* __restrict__ tw) {
    const int nb = N / R;
    const int tstep = N / (Ns * R);
    float2 v[2][R];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int j = threadIdx.x + q * 256;
        if (j < nb) {
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = buf[j + r * nb];
#pragma unroll
            for (int r = 1; r < R; r++) v[q][r] = cmulf(v[q][r], tw[r * k * tstep]);
            dft_small<R>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int j = threadIdx.x + q * 256;
        if (j < nb) {
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[base + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

__device__ inline void fft_lds(float2* buf, int N, int logN, const float2* __restrict__ tw) {
    int Ns = 1;
    for (int s = 0; s < logN / 3; s++) {
        fft_pass<8>(buf, N, Ns, tw);
        Ns *= 8;
    }
    if (logN % 3 == 2) fft_pass<4>(buf, N, Ns, tw);
    else if (logN % 3 == 1) fft_pass<2>(buf, N, Ns, tw);
}

// Rows 2 pair and 2 pair + 1 (< nrows; the first always is) of one block of
// 256 lanes, N complex values of dynamic LDS.
__device__ __forceinline__ void ramp_fft_rows(float* __restrict__ out, const float* __restrict__ in,
                                              const float* __restrict__ G, const float2* __restrict__ tw,
                                              const float* __restrict__ pw, long nrows, int nu, int N, int logN,
                                              long pair) {
    extern __shared__ float2 zbuf[];
    const long ra = 2L * pair, rb = ra + 1;
    const bool hb = rb < nrows;
    const float* xa = in + ra * nu;
    const float* xb = in + rb * nu;
    for (int m = threadIdx.x; m < N; m += 256) {
        float a = 0.f, b = 0.f;
        if (m < nu) {
            const float w = pw[m];
            a = xa[m] * w;
            if (hb) b = xb[m] * w;
        }
        zbuf[m] = make_float2(a, b);
    }
    __syncthreads();
    fft_lds(zbuf, N, logN, tw);
    for (int m = threadIdx.x; m < N; m += 256) {
        const float2 z = zbuf[m];
        const float gm = G[m];
        zbuf[m] = make_float2(z.x * gm, -z.y * gm);       // conj(Z) G
    }
    __syncthreads();
    fft_lds(zbuf, N, logN, tw);
    float* ya = out + ra * nu;
    float* yb = out + rb * nu;
    for (int j = threadIdx.x; j < nu; j += 256) {
        const float2 f = zbuf[j];
        ya[j] = f.x;
        if (hb) yb[j] = -f.y;
    }
}

__global__ __launch_bounds__(256) void ramp_fft_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                       const float* __restrict__ G, const float2* __restrict__ tw,
                                                       const float* __restrict__ pw, long nrows, int nu, int N,
                                                       int logN) {
    ramp_fft_rows(out, in, G, tw, pw, nrows, nu, N, logN, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Three-pass form for N = 1024, 2048, 4096 (nu from 257 to 2048): radix 16,
// radix 16, radix N/256.  T = N/16 lanes per complex row (one radix-16
// butterfly each; the last pass gives each lane 256/T butterflies),
// max(T, 128)/T complex rows (two detector rows each) per block.  The first pass reads the
// rows from global memory and the last pass of the inverse writes them back,
// so a row pair crosses LDS five times (the multiplication by conj and G
// happens in registers between the transforms).  LDS rows are padded by one
// complex value per 16 (index i at i + i/16), which spreads the radix-16
// writes over all banks.  Twiddles W_N^m from the hardware sine and cosine of
// m/N revolutions, powers by repeated multiplication.
// ---------------------------------------------------------------------------
__device__ inline float2 twiddle(int m, float invN) {
    const float x = -(float)m * invN;              // revolutions
    return make_float2(__builtin_amdgcn_cosf(x), __builtin_amdgcn_sinf(x));
}

// X[k] = sum_n v[n] W_16^{nk}: 4 x 4 decomposition, n = 4 n1 + n2, k = k1 + 4 k2
__device__ inline void dft16(float2* v) {
    constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, c2 = 0.70710678118654752f;
    // W_16^m = (cos, -sin)(2 pi m / 16) for m = n2 k1, n2, k1 < 4
    const float2 w[10] = {{1.f, 0.f}, {c1, -s1}, {c2, -c2}, {s1, -c1}, {0.f, -1.f}, {-s1, -c1}, {-c2, -c2},
                          {-c1, -s1}, {-1.f, 0.f}, {-c1, s1}};
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        float2 q[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
        dft_small<4>(q);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) a[n2][k1] = (n2 * k1 == 0) ? q[k1] : cmulf(q[k1], w[n2 * k1]);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft_small<4>(q);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = q[k2];
    }
}

template <int R>
__device__ inline void twiddle_run(float2* v, int m, float invN) {
    if (m == 0) return;
    const float2 w1 = twiddle(m, invN);
    float2 w = w1;
#pragma unroll
    for (int r = 1; r < R; r++) {
        v[r] = cmulf(v[r], w);
        if (r + 1 < R) w = cmulf(w, w1);
    }
}

template <int R>
__device__ inline void dft_any(float2* v) {
    if constexpr (R == 16) dft16(v);
    else dft_small<R>(v);
}

// Blocks of max(T, 128) lanes: one complex row per block from N = 2048 on
// (17 KiB of LDS, barriers over two wavefronts), and registers held to five
// wavefronts per SIMD where the LDS allows them.
template <int N>
constexpr int fft16_block() { return N / 16 < 128 ? 128 : N / 16; }

// Rows 2 pair and 2 pair + 1 (those < nrows) for the complex row of lanes
// threadIdx.x / (N / 16); every lane of the block calls it (barriers).
template <int N>
__device__ __forceinline__ void ramp_fft16_rows(float* __restrict__ out, const float* __restrict__ in,
                                                const float* __restrict__ G, const float* __restrict__ pw, long nrows,
                                                int nu, long pair) {
    constexpr int T = N / 16;            // lanes per complex row
    constexpr int RPB = fft16_block<N>() / T;   // complex rows per block
    constexpr int R3 = N / 256;          // radix of the third pass
    constexpr int B3 = 256 / T;          // third-pass butterflies per lane
    constexpr int NP = N + N / 16;       // padded LDS row
    static_assert(N == 1024 || N == 2048 || N == 4096, "three-pass plan");
    __shared__ float2 lds[RPB * NP];
    const float invN = 1.f / (float)N;
    // T >= 64: the complex row of a wavefront is uniform, so its rows are
    // addressed from scalar registers
    const int t = threadIdx.x % T, c = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    float2* buf = lds + c * NP;
    const long ra = 2L * pair, rb = ra + 1;
    const bool ha = ra < nrows, hb = rb < nrows;
    auto P = [](int i) { return i + (i >> 4); };
    float2 v[16];
    float2 u[B3][R3];

    // ---- forward transform of z = (a + i b) w ----
    // pass 1 (radix 16, Ns = 1): inputs from global memory.  nu <= N/2, so
    // inputs r >= 8 (m >= N/2) are zero and the first radix-4 stage of each
    // radix-16 butterfly has two zero inputs.
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = t + r * T;
        float a = 0.f, b = 0.f;
        if (r < 8 && m < nu) {
            const float w = pw[m];
            if (ha) a = in[ra * nu + m] * w;
            if (hb) b = in[rb * nu + m] * w;
        }
        v[r] = make_float2(a, b);
    }
    dft16(v);
#pragma unroll
    for (int r = 0; r < 16; r++) buf[P(t * 16 + r)] = v[r];
    __syncthreads();
    // pass 2 (radix 16, Ns = 16)
    {
        const int k = t & 15;
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = buf[P(t + r * T)];
        twiddle_run<16>(v, k * (N / 256), invN);
        dft16(v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) buf[P((t - k) * 16 + k + 16 * r)] = v[r];
        __syncthreads();
    }
    // pass 3 (radix R3, Ns = 256): lane holds X[j + 256 r], j = t + q T
#pragma unroll
    for (int q = 0; q < B3; q++) {
        const int j = t + q * T;
#pragma unroll
        for (int r = 0; r < R3; r++) u[q][r] = buf[P(j + 256 * r)];
        twiddle_run<R3>(u[q], j, invN);
        dft_any<R3>(u[q]);
    }
    __syncthreads();
    // conj(X) G into LDS, natural order
#pragma unroll
    for (int q = 0; q < B3; q++)
#pragma unroll
        for (int r = 0; r < R3; r++) {
            const int i = t + q * T + 256 * r;
            const float g = G[i];
            buf[P(i)] = make_float2(u[q][r].x * g, -u[q][r].y * g);
        }
    __syncthreads();
    // ---- second forward transform (the inverse, by conjugation) ----
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = buf[P(t + r * T)];
    dft16(v);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) buf[P(t * 16 + r)] = v[r];
    __syncthreads();
    {
        const int k = t & 15;
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = buf[P(t + r * T)];
        twiddle_run<16>(v, k * (N / 256), invN);
        dft16(v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) buf[P((t - k) * 16 + k + 16 * r)] = v[r];
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < B3; q++) {
        const int j = t + q * T;
#pragma unroll
        for (int r = 0; r < R3; r++) u[q][r] = buf[P(j + 256 * r)];
        twiddle_run<R3>(u[q], j, invN);
        dft_any<R3>(u[q]);
        // outputs i >= N/2 >= nu are not needed (their arithmetic is dropped)
#pragma unroll
        for (int r = 0; r < R3 / 2; r++) {
            const int i = j + 256 * r;
            if (i < nu) {
                if (ha) out[ra * nu + i] = u[q][r].x;
                if (hb) out[rb * nu + i] = -u[q][r].y;
            }
        }
    }
}

template <int N>
__global__ __launch_bounds__(fft16_block<N>()) __attribute__((amdgpu_waves_per_eu(N <= 2048 ? 5 : 1)))
void ramp_fft16_kernel(float* __restrict__ out, const float* __restrict__ in, const float* __restrict__ G,
                       const float* __restrict__ pw, long nrows, int nu) {
    constexpr int RPB = fft16_block<N>() / (N / 16);
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x / (N / 16));
    ramp_fft16_rows<N>(out, in, G, pw, nrows, nu, (long)blockIdx.x * RPB + c);
}

// ---------------------------------------------------------------------------
// Batches of uneven filter jobs (no counterpart in the reference): the row
// pairs of the jobs with one transform size N are numbered in a row (job j's
// from pair_begin), and each complex row of a block (a wavefront or more)
// finds its job and filters one pair with that job's tables.
// ---------------------------------------------------------------------------
struct RampJobDev {
    float* out;
    const float* in;
    const float* G;
    const float2* tw;
    const float* pw;
    long nrows, pair_begin;
    int nu;
};

// the last job whose first pair is <= p (jobs[0].pair_begin = 0): a binary
// search, wave-uniform, so the records are read with scalar loads
__device__ inline const RampJobDev* ramp_job_of(const RampJobDev* __restrict__ jobs, int njobs, long p) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].pair_begin <= p) lo = mid;
        else hi = mid - 1;
    }
    return jobs + lo;
}

template <int N>
__global__ __launch_bounds__(fft16_block<N>()) __attribute__((amdgpu_waves_per_eu(N <= 2048 ? 5 : 1)))
void ramp_fft16_batch_kernel(const RampJobDev* __restrict__ jobs, int njobs, long npairs) {
    constexpr int RPB = fft16_block<N>() / (N / 16);
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x / (N / 16));
    const long p = (long)blockIdx.x * RPB + c;
    const RampJobDev* J = ramp_job_of(jobs, njobs, p);
    // a complex row past the last pair filters nothing (its lanes still meet the barriers)
    ramp_fft16_rows<N>(J->out, J->in, J->G, J->pw, p < npairs ? J->nrows : 0, J->nu, p - J->pair_begin);
}

__global__ __launch_bounds__(256) void ramp_fft_batch_kernel(const RampJobDev* __restrict__ jobs, int njobs, int N,
                                                             int logN) {
    const long p = blockIdx.x;
    const RampJobDev* J = ramp_job_of(jobs, njobs, p);
    ramp_fft_rows(J->out, J->in, J->G, J->tw, J->pw, J->nrows, J->nu, N, logN, p - J->pair_begin);
}

// G = FFT(g) (real) times scale / N, and the twiddle table W_N^t, t < N, in FP64
struct RampFFTTables {
    int N = 0, logN = 0;
    std::vector<float> G;
    std::vector<float2> tw;
};

static RampFFTTables ramp_fft_tables(int nu, float da, int filter, bool equispace, float scale) {
    static std::mutex mu;
    static std::map<std::tuple<int, float, int, bool, float>, RampFFTTables> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(nu, da, filter, equispace, scale);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    RampFFTTables t;
    t.logN = 0;
    while ((1 << t.logN) < 2 * nu - 1) t.logN++;
    t.logN = std::max(t.logN, 3);
    t.N = 1 << t.logN;
    const int N = t.N;
    const std::vector<double> h = ramp_kernel_fp64(nu, (double)da, filter, equispace);
    std::vector<double> cs(N);
    for (int k = 0; k < N; k++) cs[k] = std::cos(2 * M_PI * k / N);
    // g even: G[m] = g[0] + sum_{0 < |k| < nu} h_w[nu - 1 + k] cos(2 pi k m / N)
    t.G.resize(N);
    for (int m = 0; m < N; m++) {
        double acc = h[nu - 1];
        for (int k = 1; k < nu; k++)
            acc += (h[nu - 1 + k] + h[nu - 1 - k]) * cs[(int)(((long)k * m) % N)];
        t.G[m] = (float)(acc * (double)scale / N);
    }
    t.tw.resize(N);
    for (int k = 0; k < N; k++) t.tw[k] = make_float2((float)cs[k], (float)(-std::sin(2 * M_PI * k / N)));
    if (cache.size() > 64) cache.clear();
    cache[key] = t;
    return t;
}

// Filter pre-weight per detector column: cos(gamma) for the fan beam,
// cuda/fbp.cu:60-66 (not mirrored), 1 for the parallel beam.
static std::vector<float> ramp_preweight(int nu, float da, float off_u, bool equispace) {
    std::vector<float> pw(nu);
    for (int i = 0; i < nu; i++) {
        const double ang = ((i - (nu - 1) / 2.0) - off_u) * (double)da;
        pw[i] = equispace ? 1.f : (float)std::cos(ang);
    }
    return pw;
}

// cuda/fbp.cu:97-104, cuda/fbpParallel.cu:59; the 1/L cancels cuFFT's
// unnormalised inverse
static float ramp_scale(int nview, float da, float dso, bool equispace) {
    return equispace ? (float)((double)kPI / nview * da) : (float)((double)kPI / nview * da / dso);
}

struct FilterBatchPlan::Impl {
    struct Group {
        int N, logN;
        RampJobDev* djobs;
        int njobs;
        long npairs;
    };
    std::vector<Group> groups;
    // jobs of more than 2048 columns: the single call's direct convolution,
    // one launch each
    struct Wide {
        float* out;
        const float* in;
        float* h;
        float* pw;
        long nrows;
        int nu;
        float scale;
    };
    std::vector<Wide> wide;
    std::vector<void*> mem;
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
};

// Tables of every job (the single-call filter's, see ramp_filter) uploaded
// once; jobs grouped by transform size N, each group one launch of the
// single-call kernel for that N (three-pass from N = 1024, else the generic
// one), so each job's output equals the single call's.  Jobs of more than
// 2048 columns take the single call's direct convolution, one launch each.
FilterBatchPlan::FilterBatchPlan(const std::vector<FilterBatchJobSpec>& jobs) : impl_(new Impl) {
    Impl& P = *impl_;
    std::map<int, std::vector<RampJobDev>> byN;
    for (const FilterBatchJobSpec& S : jobs) {
        if (S.nu < 1 || S.nu > 5000) fail("filter batch: detector rows of 1 to 5000 columns");
        if (2 * S.nu - 1 > 4096) {
            const std::vector<float> h = cached_ramp(S.nu, S.da, S.filter, S.equispace);
            const std::vector<float> pw = ramp_preweight(S.nu, S.da, S.off_u, S.equispace);
            Impl::Wide W{S.out, S.in, P.alloc<float>(h.size()), P.alloc<float>(S.nu), (long)S.B * S.nview * S.nv,
                         S.nu, ramp_scale(S.nview, S.da, S.dso, S.equispace)};
            CTP_CHECK(hipMemcpy(W.h, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
            CTP_CHECK(hipMemcpy(W.pw, pw.data(), sizeof(float) * S.nu, hipMemcpyHostToDevice));
            P.wide.push_back(W);
            continue;
        }
        const RampFFTTables t = ramp_fft_tables(S.nu, S.da, S.filter, S.equispace,
                                                ramp_scale(S.nview, S.da, S.dso, S.equispace));
        const std::vector<float> pw = ramp_preweight(S.nu, S.da, S.off_u, S.equispace);
        float* dG = P.alloc<float>(t.N);
        float2* dtw = P.alloc<float2>(t.N);
        float* dpw = P.alloc<float>(S.nu);
        CTP_CHECK(hipMemcpy(dG, t.G.data(), sizeof(float) * t.N, hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(dtw, t.tw.data(), sizeof(float2) * t.N, hipMemcpyHostToDevice));
        CTP_CHECK(hipMemcpy(dpw, pw.data(), sizeof(float) * S.nu, hipMemcpyHostToDevice));
        RampJobDev J{};
        J.out = S.out;
        J.in = S.in;
        J.G = dG;
        J.tw = dtw;
        J.pw = dpw;
        J.nrows = (long)S.B * S.nview * S.nv;
        J.nu = S.nu;
        byN[t.N].push_back(J);
    }
    for (auto& kv : byN) {
        Impl::Group G{kv.first, 0, nullptr, (int)kv.second.size(), 0};
        while ((1 << G.logN) < G.N) G.logN++;
        for (RampJobDev& J : kv.second) {
            J.pair_begin = G.npairs;
            G.npairs += ceil_div(J.nrows, 2L);
        }
        if (G.npairs > 0x7fffffffL) fail("filter batch: too many detector rows for one batch; split the jobs");
        G.djobs = P.alloc<RampJobDev>(kv.second.size());
        CTP_CHECK(hipMemcpy(G.djobs, kv.second.data(), sizeof(RampJobDev) * kv.second.size(), hipMemcpyHostToDevice));
        P.groups.push_back(G);
    }
}

FilterBatchPlan::~FilterBatchPlan() = default;

int FilterBatchPlan::launches() const { return (int)(impl_->groups.size() + impl_->wide.size()); }

void FilterBatchPlan::run(hipStream_t s) {
    TimedLaunch tl("ramp_filter_batch", s);
    for (const Impl::Group& G : impl_->groups) {
        if (G.N >= 1024) {
            auto launch = [&](auto kernel, int block) {
                const int rpb = block / (G.N / 16);
                kernel<<<(unsigned)ceil_div(G.npairs, (long)rpb), block, 0, s>>>(G.djobs, G.njobs, G.npairs);
            };
            if (G.N == 1024) launch(ramp_fft16_batch_kernel<1024>, fft16_block<1024>());
            else if (G.N == 2048) launch(ramp_fft16_batch_kernel<2048>, fft16_block<2048>());
            else launch(ramp_fft16_batch_kernel<4096>, fft16_block<4096>());
        } else {
            ramp_fft_batch_kernel<<<(unsigned)G.npairs, 256, sizeof(float2) * G.N, s>>>(G.djobs, G.njobs, G.N,
                                                                                       G.logN);
        }
        CTP_LAUNCH_CHECK();
    }
    for (const Impl::Wide& W : impl_->wide) {
        // the single call's launch shapes (ramp_filter)
        const size_t sh4 = sizeof(float) * ((2 * W.nu - 1 + 4) + 4 * W.nu);
        if (sh4 <= 65536) {
            ramp_conv_kernel<4><<<(unsigned)ceil_div(W.nrows, 4L), 256, sh4, s>>>(W.out, W.in, W.h, W.pw, W.nrows,
                                                                                 W.nu, W.scale, nullptr, 1, 1);
        } else {
            const size_t sh1 = sizeof(float) * ((2 * W.nu - 1 + 4) + W.nu);
            ramp_conv_kernel<1><<<(unsigned)W.nrows, 256, sh1, s>>>(W.out, W.in, W.h, W.pw, W.nrows, W.nu, W.scale,
                                                                    nullptr, 1, 1);
        }
        CTP_LAUNCH_CHECK();
    }
}

void ramp_filter(float* out, const float* prj, int B, int nu, int nv, int nview, float da, float off_u, float dsd,
                 float dso, int filter, bool equispace, hipStream_t stream) {
    (void)dsd;
    if (nu <= 0) return;
    if (nu > 5000) fail("ramp filter: nu > 5000 is not supported");
    const std::vector<float> pw = ramp_preweight(nu, da, off_u, equispace);
    const float scale = ramp_scale(nview, da, dso, equispace);
    const long nrows = (long)B * nview * nv;
    // ramp_mode: 0 = FFT convolution where N = 2^p >= 2nu - 1 is at most 4096
    // (three-pass kernel for N >= 1024), 1 = direct, 2 = the generic FFT kernel
    const int mode = tuning_get("ramp_mode", 0);
    if (mode < 0 || mode > 2) fail("ramp_mode must be 0 (FFT), 1 (direct) or 2 (generic FFT)");
    if (mode != 1 && 2 * nu - 1 <= 4096) {
        const RampFFTTables t = ramp_fft_tables(nu, da, filter, equispace, scale);
        Workspace& ws = workspace_for(stream);
        Plan pl;
        pl.add<float>(t.N);
        pl.add<float2>(t.N);
        pl.add<float>(nu);
        ws.reserve(pl.bytes);
        float* dG = ws.take_n<float>(t.N);
        float2* dtw = ws.take_n<float2>(t.N);
        float* dpw = ws.take_n<float>(nu);
        CTP_CHECK(hipMemcpyAsync(dG, t.G.data(), sizeof(float) * t.N, hipMemcpyHostToDevice, stream));
        CTP_CHECK(hipMemcpyAsync(dtw, t.tw.data(), sizeof(float2) * t.N, hipMemcpyHostToDevice, stream));
        CTP_CHECK(hipMemcpyAsync(dpw, pw.data(), sizeof(float) * nu, hipMemcpyHostToDevice, stream));
        {
            TimedLaunch tl("ramp_filter", stream);
            const long pairs = ceil_div(nrows, 2);
            if (mode == 0 && t.N >= 1024) {
                auto launch = [&](auto kernel, int block) {
                    const int rpb = block / (t.N / 16);
                    kernel<<<(unsigned)ceil_div(pairs, (long)rpb), block, 0, stream>>>(out, prj, dG, dpw, nrows, nu);
                };
                if (t.N == 1024) launch(ramp_fft16_kernel<1024>, fft16_block<1024>());
                else if (t.N == 2048) launch(ramp_fft16_kernel<2048>, fft16_block<2048>());
                else launch(ramp_fft16_kernel<4096>, fft16_block<4096>());
            } else {
                ramp_fft_kernel<<<(unsigned)pairs, 256, sizeof(float2) * t.N, stream>>>(out, prj, dG, dtw, dpw, nrows,
                                                                                      nu, t.N, t.logN);
            }
            CTP_LAUNCH_CHECK();
        }
        CTP_CHECK(hipStreamSynchronize(stream));
        return;
    }
    const std::vector<float> h = cached_ramp(nu, da, filter, equispace);
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>(2 * nu - 1);
    pl.add<float>(nu);
    ws.reserve(pl.bytes);
    float* dh = ws.take_n<float>(2 * nu - 1);
    float* dpw = ws.take_n<float>(nu);
    CTP_CHECK(hipMemcpyAsync(dh, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dpw, pw.data(), sizeof(float) * nu, hipMemcpyHostToDevice, stream));
    // LDS: the kernel (2nu+3 floats) and RB input rows; stays within 64 KB
    const size_t sh4 = sizeof(float) * ((2 * nu - 1 + 4) + 4 * nu);
    {
    TimedLaunch t("ramp_filter", stream);
    if (sh4 <= 65536) {
        ramp_conv_kernel<4><<<(unsigned)ceil_div(nrows, 4), 256, sh4, stream>>>(out, prj, dh, dpw, nrows, nu, scale,
                                                                             nullptr, 1, 1);
    } else {
        const size_t sh1 = sizeof(float) * ((2 * nu - 1 + 4) + nu);
        ramp_conv_kernel<1><<<(unsigned)nrows, 256, sh1, stream>>>(out, prj, dh, dpw, nrows, nu, scale, nullptr, 1,
                                                                   1);
    }
    CTP_LAUNCH_CHECK();
    }
    CTP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// tomosynthesis FDK filter, cuda/fbpTomo.cu:47-368
// ---------------------------------------------------------------------------
// Per view, the reference builds the equispace Ram-Lak kernel for the virtual
// pitch vdu = du cos(b) (b: the view's tilt in the x-z plane), multiplies its
// DFT by a window W(f) of the x and z frequencies (GetFrequency, :47-65: the
// frequency is clipped to the cutoff from above only, so W is not symmetric),
// and keeps the real part of the inverse DFT.  For a real kernel that real
// part is the inverse DFT with the symmetrised window (W(f) + W(-f)) / 2, i.e.
// the circular convolution of the Ram-Lak kernel with the real, even kernel
// w_s; it is computed here in FP64 and applied as an exact linear convolution
// read at nu - 1 (as the padded FFT filtering, :330-356), scaled by
// pi / nview * vdu.
static double tomo_freq(int i, int L, double vdu, double d, double weight, double cutoff) {
    const double fu = (i > (L + 1) / 2) ? (i - L) / (double)L / vdu : i / (double)L / vdu;
    double w = fu * weight * d * 2;
    if (w > cutoff) w = cutoff;
    return w;
}

static std::vector<double> tomo_kernel(int nu, double vdu, double cosb, double sinb, double dx, double dz, int filter,
                                       double cutx, double cutz) {
    const int L = 2 * nu - 1;
    const double PI = (double)kPI;
    std::vector<double> r(L, 0.0), W(L, 1.0), ws(L), h(L, 0.0), ct(L);
    for (int i = 0; i < L; i++) {
        const int k = i - (nu - 1);
        if (k == 0) r[i] = 1.0 / (4 * vdu * vdu);
        else if (k % 2 != 0) r[i] = -1.0 / (PI * PI * (double)k * k * vdu * vdu);
    }
    if (filter == FILTER_HAMMING || filter == FILTER_HANN) {
        const double A = filter == FILTER_HAMMING ? 0.54 : 0.5, Bw = filter == FILTER_HAMMING ? 0.46 : 0.5;
        for (int i = 0; i < L; i++) {
            const double wx = tomo_freq(i, L, vdu, dx, cosb, cutx), wz = tomo_freq(i, L, vdu, dz, sinb, cutz);
            W[i] = (A + Bw * std::cos(PI * wx / cutx)) * (A + Bw * std::cos(PI * wz / cutz));
        }
    }
    for (int m = 0; m < L; m++) ct[m] = std::cos(2 * M_PI * m / L);
    for (int k = 0; k < L; k++) {           // w_s = IDFT of the symmetrised window (real, even)
        double acc = 0;
        for (int f = 0; f < L; f++) acc += 0.5 * (W[f] + W[(L - f) % L]) * ct[(long)f * k % L];
        ws[k] = acc / L;
    }
    for (int i = 0; i < L; i++) {           // circular convolution r * w_s
        double acc = 0;
        for (int j = 0; j < L; j++) acc += r[j] * ws[((i - j) % L + L) % L];
        h[i] = acc;
    }
    return h;
}

void tomo_filter(float* out, const float* prj, const float* det_center_host, const float* src_host, int B, int nu,
                 int nv, int nview, float du, float dx, float dz, int filter, float cutx, float cutz,
                 hipStream_t stream) {
    if (nu <= 0 || nv <= 0 || nview <= 0 || B <= 0) return;
    if (nu > 5000) fail("tomo filter: nu > 5000 is not supported");
    if ((filter == FILTER_HAMMING || filter == FILTER_HANN) && !(cutx > 0 && cutz > 0))
        fail("tomo filter: cutoffs must be positive");
    const int L = 2 * nu - 1;
    std::vector<float> h((size_t)nview * L), sc(nview), pw(nu, 1.f);
    for (int v = 0; v < nview; v++) {
        const double ex = (double)det_center_host[3 * v] - src_host[3 * v];
        const double ez = (double)det_center_host[3 * v + 2] - src_host[3 * v + 2];
        const double r = std::sqrt(ex * ex + ez * ez);
        if (!(r > 0)) fail("tomo filter: source and detector centre coincide in x-z in view " + std::to_string(v));
        const double sinb = ex / r, cosb = -ez / r;           // cuda/fbpTomo.cu:130-133
        const double vdu = (double)du * cosb;
        if (vdu == 0) fail("tomo filter: view " + std::to_string(v) + " is parallel to the detector");
        const std::vector<double> hk = tomo_kernel(nu, vdu, cosb, sinb, dx, dz, filter, cutx, cutz);
        for (int i = 0; i < L; i++) h[(size_t)v * L + i] = (float)hk[i];
        sc[v] = (float)((double)kPI / nview * vdu);
    }
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>((size_t)nview * L);
    pl.add<float>(nview);
    pl.add<float>(nu);
    ws.reserve(pl.bytes);
    float* dh = ws.take_n<float>((size_t)nview * L);
    float* dsc = ws.take_n<float>(nview);
    float* dpw = ws.take_n<float>(nu);
    CTP_CHECK(hipMemcpyAsync(dh, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dsc, sc.data(), sizeof(float) * nview, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dpw, pw.data(), sizeof(float) * nu, hipMemcpyHostToDevice, stream));
    const long nrows = (long)B * nview * nv;
    const size_t sh1 = sizeof(float) * ((2 * nu - 1 + 4) + nu);
    {
        TimedLaunch t("tomo_filter", stream);
        ramp_conv_kernel<1><<<(unsigned)nrows, 256, sh1, stream>>>(out, prj, dh, dpw, nrows, nu, 1.f, dsc, nv, nview);
        CTP_LAUNCH_CHECK();
    }
    CTP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// fan FBP backprojection
// ---------------------------------------------------------------------------
struct FBPView {
    float c, s;
    float pad0, pad1;
};

// Detector position pu (cells) and distance weight of the pixel centre (px,
// py) in the view of direction (c, s), cuda/fbp.cu:288-321 (parallel:
// cuda/fbpParallel.cu:224-236, weight 1).  Contraction is off and the FMAs
// are explicit, so that every kernel below gets the same bits wherever it is
// inlined.
template <bool FAN>
__device__ inline void fbp_view_point(float px, float py, float c, float s, float du, float off_u, float half,
                                      float dso, float& pu, float& dist) {
#pragma clang fp contract(off)
    const float rx = fmaf(px, c, py * s);
    const float ry = fmaf(py, c, -(px * s));
    float a = rx;
    dist = 1.f;
    if (FAN) {
        const float ys = dso + ry;
        a = atanf(fast_div(rx, ys));
        dist = fast_div(dso * dso, fmaf(rx, rx, ys * ys));
    }
    pu = half - (fast_div(a, du) + off_u);
}

// FAN: bpFanKernel3D (cuda/fbp.cu:264-359); otherwise bpParallelKernel3D
// (cuda/fbpParallel.cu:200-255): u = x cos + y sin, no distance weight.
template <int CH, bool FAN>
__global__ __launch_bounds__(256) void fan_fbp_bp_kernel(float* __restrict__ img, float* __restrict__ part,
                                                         const float* __restrict__ q, const FBPView* __restrict__ views,
                                                         int C, int cs, int nchunk, int vpg, Grid g, Det d, int nview,
                                                         float dso) {
    const int ix = blockIdx.x * 16 + (threadIdx.x & 15);
    const int iy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int grp = blockIdx.z / nchunk;
    const int c0 = (blockIdx.z - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;
    // pixel centre, cuda/fbp.cu:288
    const float px = ((float)ix + 0.5f - g.nx / 2.f) * g.dx + g.cx;
    const float py = ((float)iy + 0.5f - g.ny / 2.f) * g.dy + g.cy;
    float acc[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) acc[j] = 0.f;
    const float half = (d.nu - 1.0f) / 2.0f;
    const size_t vstride = (size_t)d.nu * cs;
    for (int view = vbeg; view < vend; view++) {
        const FBPView V = views[view];
        float pu, dist;
        fbp_view_point<FAN>(px, py, V.c, V.s, d.du, d.off_u, half, dso, pu, dist);
        // InterpolateXY with truncation, cuda/cudaMath.cu:227-260
        if (pu < 0.f || pu >= (float)d.nu) continue;
        const int u0 = (int)pu;
        const int u1 = min(u0 + 1, d.nu - 1);
        const float wu = pu - (float)u0;
        const float w0 = (1.f - wu) * dist, w1 = wu * dist;
        const float* __restrict__ p0 = q + view * vstride + (size_t)u0 * cs + c0;
        const float* __restrict__ p1 = q + view * vstride + (size_t)u1 * cs + c0;
        if constexpr (CH % 4 == 0) {
#pragma unroll
            for (int j = 0; j < CH; j += 4) {
                const float4 a0 = *reinterpret_cast<const float4*>(p0 + j);
                const float4 a1 = *reinterpret_cast<const float4*>(p1 + j);
                acc[j] = fmaf(w0, a0.x, fmaf(w1, a1.x, acc[j]));
                acc[j + 1] = fmaf(w0, a0.y, fmaf(w1, a1.y, acc[j + 1]));
                acc[j + 2] = fmaf(w0, a0.z, fmaf(w1, a1.z, acc[j + 2]));
                acc[j + 3] = fmaf(w0, a0.w, fmaf(w1, a1.w, acc[j + 3]));
            }
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) acc[j] = fmaf(w0, p0[j], fmaf(w1, p1[j], acc[j]));
        }
    }
    store_bp_result<CH>(img, part, acc, grp, c0, C, (size_t)g.nx * g.ny, (size_t)iy * g.nx + ix);
}

// The backprojection above with LPC lanes per pixel: lane s of a pixel holds
// the float4s at channel offsets 4 (s + LPC k), so each load instruction reads
// a contiguous 16 LPC-byte piece of a detector cell and a wavefront's
// 64 / LPC pixels (a WX x WY tile) touch few cells.  The per-view geometry is
// spread over the lanes of a pixel: lane s evaluates view v0 + s, and the
// pixel's lanes read each view's detector position and distance weight from
// it.  Same arithmetic per channel as fan_fbp_bp_kernel.
template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void fan_fbp_bp_split_kernel(float* __restrict__ img, float* __restrict__ part,
                                                               const float* __restrict__ q,
                                                               const FBPView* __restrict__ views, int C, int cs,
                                                               int nchunk, int vpg, Grid g, Det d, int nview,
                                                               float dso) {
    static_assert(CH % (4 * LPC) == 0 && (LPC == 2 || LPC == 4), "a lane holds whole float4s");
    constexpr int NQ = CH / (4 * LPC);
    constexpr int WX = LPC == 2 ? 8 : 4, WY = 64 / LPC / WX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = lane / LPC, sub = lane % LPC, lead = lane & ~(LPC - 1);
    const int ix = blockIdx.x * (2 * WX) + (wave & 1) * WX + p % WX;
    const int iy = blockIdx.y * (2 * WY) + (wave >> 1) * WY + p / WX;
    const int grp = blockIdx.z / nchunk;
    const int c0 = (blockIdx.z - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    if (ix >= g.nx || iy >= g.ny) return;   // the lanes of a pixel leave together
    // pixel centre, cuda/fbp.cu:288
    const float px = ((float)ix + 0.5f - g.nx / 2.f) * g.dx + g.cx;
    const float py = ((float)iy + 0.5f - g.ny / 2.f) * g.dy + g.cy;
    float acc[4 * NQ];
#pragma unroll
    for (int j = 0; j < 4 * NQ; j++) acc[j] = 0.f;
    const float half = (d.nu - 1.0f) / 2.0f;
    const size_t vstride = (size_t)d.nu * cs;
    const unsigned lane_b = 16u * sub;   // the lane's first float of a cell, in bytes
    for (int v0 = vbeg; v0 < vend; v0 += LPC) {
        float mpu = -1.f, mdist = 0.f;
        if (v0 + sub < vend) {
            const FBPView V = views[v0 + sub];
            fbp_view_point<FAN>(px, py, V.c, V.s, d.du, d.off_u, half, dso, mpu, mdist);
        }
#pragma unroll
        for (int t = 0; t < LPC; t++) {
            const int view = v0 + t;
            if (view >= vend) break;   // uniform
            const float pu = group_bcast<LPC>(mpu, t, lead);     // DPP within the pixel's lanes
            const float dist = group_bcast<LPC>(mdist, t, lead);
            // InterpolateXY with truncation, cuda/cudaMath.cu:227-260
            if (pu < 0.f || pu >= (float)d.nu) continue;
            const int u0 = (int)pu;
            const int u1 = min(u0 + 1, d.nu - 1);
            const float wu = pu - (float)u0;
            const float w0 = (1.f - wu) * dist, w1 = wu * dist;
            // a view's cells from one scalar base plus 32-bit per-lane byte offsets
            const char* __restrict__ vb = reinterpret_cast<const char*>(q + c0 + view * vstride);
            const unsigned o0 = (unsigned)u0 * (unsigned)cs * 4u + lane_b;
            const unsigned o1 = (unsigned)u1 * (unsigned)cs * 4u + lane_b;
#pragma unroll
            for (int k = 0; k < NQ; k++) {
                const float4 a0 = *reinterpret_cast<const float4*>(vb + o0 + 16u * LPC * k);
                const float4 a1 = *reinterpret_cast<const float4*>(vb + o1 + 16u * LPC * k);
                acc[4 * k] = fmaf(w0, a0.x, fmaf(w1, a1.x, acc[4 * k]));
                acc[4 * k + 1] = fmaf(w0, a0.y, fmaf(w1, a1.y, acc[4 * k + 1]));
                acc[4 * k + 2] = fmaf(w0, a0.z, fmaf(w1, a1.z, acc[4 * k + 2]));
                acc[4 * k + 3] = fmaf(w0, a0.w, fmaf(w1, a1.w, acc[4 * k + 3]));
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

// The split backprojection with each wavefront's detector cells staged in
// LDS.  Per group of LPC views the wavefront's pixels reduce their cells to
// one range [min u0, max u1] per view (a reduction over the lanes that hold
// the same view).  Each view's range is copied into the wavefront's 4 KiB of
// LDS by whole LDS DMA instructions (global_load_lds_dwordx4: lane l copies
// float4 l mod Q of cell l / Q, i.e. 64 / Q cells per instruction; Q = CH / 4
// float4 per cell), with a uniform base per instruction and a per-lane offset
// that is the same for all of them; the pixels then read their two cells from
// LDS.  In the direct form every pixel of a wavefront load looks the cells up
// in the vector L1 again (64 / LPC lookups of a few cells); here each cell
// passes the L1 once per view.  A group needing more than four instructions
// reads global memory as the split kernel does.  The view table is read with
// uniform addresses (scalar loads).  A 16-byte LDS read serves groups of
// lanes, i.e. groups of pixels that mostly read different cells: pixel p
// takes its float4s k = 0 .. NQ-1 in the order k ^ (p mod NQ), so that
// neighbouring pixels read different parts of their cells, i.e. different
// banks; a lane's accumulator k holds the channels of float4 k ^ (p mod NQ)
// (measured at 512^2 x 32: other orders within 2%, the kernel is bound by
// instruction issue, not by LDS).  Every channel sees the arithmetic of
// fan_fbp_bp_split_kernel in the same order: the same results.  Needs
// nu < 2^30 and nu >= four instructions' cells (checked by the host).
template <int CH, int LPC, bool FAN>
__global__ __launch_bounds__(256) void fan_fbp_bp_stage_kernel(float* __restrict__ img, float* __restrict__ part,
                                                               const float* __restrict__ q,
                                                               const FBPView* __restrict__ views, int C, int cs,
                                                               int nchunk, int vpg, Grid g, Det d, int nview,
                                                               float dso) {
    static_assert(CH % (4 * LPC) == 0 && (LPC == 2 || LPC == 4), "a lane holds whole float4s");
    constexpr int NQ = CH / (4 * LPC);
    static_assert(NQ == 1 || NQ == 2 || NQ == 4, "float4 order by pixel");
    constexpr int WX = LPC == 2 ? 8 : 4, WY = 64 / LPC / WX;
    constexpr int Q = CH / 4;               // float4 per staged cell
    constexpr int CPI = 64 / Q;             // cells per copy instruction
    constexpr int KI = 4;                   // copy instructions (KiB) per wavefront and group
    __shared__ float4 stage[4][64 * KI];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = lane / LPC, sub = lane % LPC, lead = lane & ~(LPC - 1);
    const int sw = p & (NQ - 1);            // float4 order of this pixel
    const int ix = blockIdx.x * (2 * WX) + (wave & 1) * WX + p % WX;
    const int iy = blockIdx.y * (2 * WY) + (wave >> 1) * WY + p / WX;
    const int grp = blockIdx.z / nchunk;
    const int c0 = (blockIdx.z - grp * nchunk) * CH;
    const int vbeg = grp * vpg, vend = min(nview, vbeg + vpg);
    // every lane of the wavefront takes part in the copies and reductions;
    // lanes of pixels outside the image add nothing and store nothing
    const bool valid = ix < g.nx && iy < g.ny;
    const float px = ((float)min(ix, g.nx - 1) + 0.5f - g.nx / 2.f) * g.dx + g.cx;   // cuda/fbp.cu:288
    const float py = ((float)min(iy, g.ny - 1) + 0.5f - g.ny / 2.f) * g.dy + g.cy;
    float acc[4 * NQ];
#pragma unroll
    for (int j = 0; j < 4 * NQ; j++) acc[j] = 0.f;
    const float half = (d.nu - 1.0f) / 2.0f;
    const size_t vstride = (size_t)d.nu * cs;
    int fo[NQ];                             // float4 index within a cell of accumulator k
#pragma unroll
    for (int k = 0; k < NQ; k++) fo[k] = sub + LPC * (k ^ sw);
    const unsigned voff = (unsigned)(lane / Q) * (unsigned)cs * 4u + (unsigned)(lane % Q) * 16u;   // copy offset
    float4* buf = stage[wave];
    for (int v0 = vbeg; v0 < vend; v0 += LPC) {
        // detector position and weight of view v0 + sub
        float vc = 0.f, vs = 0.f;
#pragma unroll
        for (int t = 0; t < LPC; t++) {
            const FBPView V = views[min(v0 + t, nview - 1)];   // uniform address
            if (sub == t) { vc = V.c; vs = V.s; }
        }
        float mpu = -1.f, mdist = 0.f;
        int lo = 1 << 30, hi = -1;   // empty
        if (v0 + sub < vend) {
            fbp_view_point<FAN>(px, py, vc, vs, d.du, d.off_u, half, dso, mpu, mdist);
            if (valid && mpu >= 0.f && mpu < (float)d.nu) {
                lo = (int)mpu;
                hi = min(lo + 1, d.nu - 1);
            }
        }
        lo = class_reduce<LPC, false>(lo);
        hi = class_reduce<LPC, true>(hi);
        // view t: copy instructions [slot[t], slot[t+1]) of cells from sbeg[t]
        // (moved back where the instructions would pass the detector's end)
        int sbeg[LPC], slot[LPC + 1];
        slot[0] = 0;
#pragma unroll
        for (int t = 0; t < LPC; t++) {
            const int s0 = __builtin_amdgcn_readlane(lo, t);
            const int n = max(__builtin_amdgcn_readlane(hi, t) - s0 + 1, 0);
            const int m = (n + CPI - 1) / CPI;
            sbeg[t] = min(s0, max(d.nu - m * CPI, 0));
            slot[t + 1] = slot[t] + m;
        }
        const bool staged = slot[LPC] <= KI;   // uniform
        if (staged) {
            // the previous group's reads of this LDS are complete
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            // nu >= KI * CPI (host): every copied cell lies in its detector row
            const char* vb0 = reinterpret_cast<const char*>(q + c0 + (size_t)v0 * vstride) + voff;
#pragma unroll
            for (int t = 0; t < LPC; t++) {
                const char* vbt = vb0 + (size_t)t * vstride * 4 + (unsigned)sbeg[t] * (unsigned)cs * 4u;
                for (int i = slot[t]; i < slot[t + 1]; i++)
                    __builtin_amdgcn_global_load_lds(
                            (const void*)(vbt + (unsigned)((i - slot[t]) * CPI) * (unsigned)cs * 4u),
                            (__attribute__((address_space(3))) void*)(buf + 64 * i), 16, 0, 0);
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __builtin_amdgcn_wave_barrier();
        }
        // the two forms are separate loops, so that each keeps its own
        // address space (LDS reads, global loads)
        auto view_step = [&](int t, auto&& load) {
            const float pu = group_bcast<LPC>(mpu, t, lead);
            const float dist = group_bcast<LPC>(mdist, t, lead);
            // InterpolateXY with truncation, cuda/cudaMath.cu:227-260
            if (!valid || pu < 0.f || pu >= (float)d.nu) return;
            const int u0 = (int)pu;
            const int u1 = min(u0 + 1, d.nu - 1);
            const float wu = pu - (float)u0;
            const float w0 = (1.f - wu) * dist, w1 = wu * dist;
            float4 a0[NQ], a1[NQ];
            load(u0, u1, a0, a1);
#pragma unroll
            for (int k = 0; k < NQ; k++) {
                acc[4 * k] = fmaf(w0, a0[k].x, fmaf(w1, a1[k].x, acc[4 * k]));
                acc[4 * k + 1] = fmaf(w0, a0[k].y, fmaf(w1, a1[k].y, acc[4 * k + 1]));
                acc[4 * k + 2] = fmaf(w0, a0[k].z, fmaf(w1, a1[k].z, acc[4 * k + 2]));
                acc[4 * k + 3] = fmaf(w0, a0[k].w, fmaf(w1, a1[k].w, acc[4 * k + 3]));
            }
        };
        if (staged) {
#pragma unroll
            for (int t = 0; t < LPC; t++) {
                if (v0 + t >= vend) break;   // uniform
                const float4* vb = buf + (64 * slot[t] - Q * sbeg[t]);
                view_step(t, [&](int u0, int u1, float4* a0, float4* a1) {
#pragma unroll
                    for (int k = 0; k < NQ; k++) {
                        a0[k] = vb[u0 * Q + fo[k]];
                        a1[k] = vb[u1 * Q + fo[k]];
                    }
                });
            }
        } else {
#pragma unroll
            for (int t = 0; t < LPC; t++) {
                const int view = v0 + t;
                if (view >= vend) break;   // uniform
                const float4* __restrict__ vb = reinterpret_cast<const float4*>(q + c0 + view * vstride);
                view_step(t, [&](int u0, int u1, float4* a0, float4* a1) {
                    const unsigned o0 = (unsigned)u0 * (unsigned)(cs / 4), o1 = (unsigned)u1 * (unsigned)(cs / 4);
#pragma unroll
                    for (int k = 0; k < NQ; k++) {
                        a0[k] = vb[o0 + fo[k]];
                        a1[k] = vb[o1 + fo[k]];
                    }
                });
            }
        }
    }
    if (!valid) return;
    const size_t npix = (size_t)g.nx * g.ny, pix = (size_t)iy * g.nx + ix;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int cc = c0 + 4 * fo[k] + e;
            if (cc >= C) continue;
            if (part) part[((size_t)grp * C + cc) * npix + pix] = acc[4 * k + e];
            else img[(size_t)cc * npix + pix] += acc[4 * k + e];
        }
    }
}

void fan_fbp_bp(float* img, const float* prj, const float* deg_host, int B, const Grid& g, const Det& d, int nview,
                float dsd, float dso, hipStream_t stream, bool fan) {
    (void)dsd;
    const int C = B * g.nz;
    int CH = 1;
    while (CH < C && CH < 32) CH *= 2;
    const int nchunk = ceil_div(C, CH);
    const int cs = nchunk * CH;
    std::vector<FBPView> views(nview);
    for (int i = 0; i < nview; i++) {
        views[i].c = (float)std::cos((double)deg_host[i]);
        views[i].s = (float)std::sin((double)deg_host[i]);
        views[i].pad0 = views[i].pad1 = 0;
    }
    // detector row coordinate of each slice centre, cuda/fbp.cu:288,321
    std::vector<SliceRow> srow(g.nz);
    for (int z = 0; z < g.nz; z++) {
        const double zc = (z + 0.5 - g.nz / 2.0) * g.dz + g.cz;
        const double pv = zc / d.dv + d.off_v + (d.nv - 1.0) / 2.0;
        SliceRow r{-1, -1, 0.f, 0};
        if (pv >= 0 && pv < d.nv) {
            const int r0 = (int)pv;
            r = SliceRow{r0, std::min(r0 + 1, d.nv - 1), (float)(pv - r0), 0};
        }
        srow[z] = r;
    }
    const size_t npix = (size_t)g.nx * g.ny;
    const int groups = bp_view_groups(npix, nchunk, nview);
    const int vpg = ceil_div(nview, groups);
    Workspace& ws = workspace_for(stream);
    Plan pl;
    pl.add<float>((size_t)nview * d.nu * cs);
    pl.add<FBPView>(nview);
    pl.add<SliceRow>(g.nz);
    if (groups > 1) pl.add<float>((size_t)groups * C * npix);
    ws.reserve(pl.bytes);
    float* q = ws.take_n<float>((size_t)nview * d.nu * cs);
    FBPView* dviews = ws.take_n<FBPView>(nview);
    SliceRow* dsrow = ws.take_n<SliceRow>(g.nz);
    float* part = groups > 1 ? ws.take_n<float>((size_t)groups * C * npix) : nullptr;
    CTP_CHECK(hipMemcpyAsync(dviews, views.data(), sizeof(FBPView) * nview, hipMemcpyHostToDevice, stream));
    CTP_CHECK(hipMemcpyAsync(dsrow, srow.data(), sizeof(SliceRow) * g.nz, hipMemcpyHostToDevice, stream));
    slices_to_channel_last(q, prj, dsrow, C, cs, g.nz, d, nview, stream);
    int lpc = tuning_get("fbp_bp_lanes_per_pixel", 4);
    if (lpc == 4 && CH % 16 != 0) lpc = 2;
    if ((lpc == 2 || lpc == 4) && CH >= 8 && CH % (4 * lpc) == 0) {
        dim3 sgrid(ceil_div(g.nx, lpc == 2 ? 16 : 8), ceil_div(g.ny, 8), nchunk * groups), block(256);
        // staged cells (measured on MI355X, fan 256^2 to 512^2 x 32 slices:
        // 0.93 to 0.96 of the split kernel's time; slower at 8 and 16
        // channels, where a view's copy and range reduction serve fewer
        // multiply-adds).  Whole copy instructions of 64 / (CH / 4) cells, at
        // most four per view group, must fit in a detector row.
        const int stage_mode = tuning_get("fbp_bp_stage", 0);
        const bool stage = (stage_mode == 1 || (stage_mode == 0 && CH == 32 && lpc == 4)) && d.nu < (1 << 30) &&
                           d.nu >= 4 * (256 / CH);
        {
        TimedLaunch t(fan ? "fan_fbp_bp" : "par_pixel_bp", stream);
#define CTP_FBPS_LAUNCH(N, L)                                                                                 \
    if (CH == N && lpc == L) {                                                                                \
        if (stage) {                                                                                           \
            if (fan) fan_fbp_bp_stage_kernel<N, L, true><<<sgrid, block, 0, stream>>>(img, part, q, dviews, C,  \
                                                                                  cs, nchunk, vpg, g, d, nview, dso); \
            else fan_fbp_bp_stage_kernel<N, L, false><<<sgrid, block, 0, stream>>>(img, part, q, dviews, C,    \
                                                                                   cs, nchunk, vpg, g, d, nview, dso); \
        } else if (fan) fan_fbp_bp_split_kernel<N, L, true><<<sgrid, block, 0, stream>>>(img, part, q, dviews, C, \
                                                                                     cs, nchunk, vpg, g, d, nview, dso); \
        else fan_fbp_bp_split_kernel<N, L, false><<<sgrid, block, 0, stream>>>(img, part, q, dviews, C, cs,    \
                                                                               nchunk, vpg, g, d, nview, dso);  \
    }
        CTP_FBPS_LAUNCH(8, 2) CTP_FBPS_LAUNCH(16, 2) CTP_FBPS_LAUNCH(16, 4) CTP_FBPS_LAUNCH(32, 2)
        CTP_FBPS_LAUNCH(32, 4)
#undef CTP_FBPS_LAUNCH
        CTP_LAUNCH_CHECK();
        if (part) reduce_parts(img, part, groups, (size_t)C * npix, stream);
        }
        CTP_CHECK(hipStreamSynchronize(stream));
        return;
    }
    dim3 grid(ceil_div(g.nx, 16), ceil_div(g.ny, 16), nchunk * groups), block(256);
    {
    TimedLaunch t(fan ? "fan_fbp_bp" : "par_pixel_bp", stream);
#define CTP_FBP_CASE(N)                                                                                        \
    case N:                                                                                                    \
        if (fan) fan_fbp_bp_kernel<N, true><<<grid, block, 0, stream>>>(img, part, q, dviews, C, cs, nchunk, vpg, g, \
                                                                        d, nview, dso);                           \
        else fan_fbp_bp_kernel<N, false><<<grid, block, 0, stream>>>(img, part, q, dviews, C, cs, nchunk, vpg, g, d, \
                                                                     nview, dso);                                  \
        break;
    switch (CH) {
        CTP_FBP_CASE(1) CTP_FBP_CASE(2) CTP_FBP_CASE(4) CTP_FBP_CASE(8) CTP_FBP_CASE(16) CTP_FBP_CASE(32)
        default: fail("fbp bp: bad channel chunk");
    }
#undef CTP_FBP_CASE
    CTP_LAUNCH_CHECK();
    if (part) reduce_parts(img, part, groups, (size_t)C * npix, stream);
    }
    CTP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace ctp
