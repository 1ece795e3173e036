// Stage (c): inverse Fourier readout of the order >=2 spectrum into time bins.
//
// The reference reads out each active OM by demodulating its spectrum by the
// OM time origin, interpolating the 65-node table grid onto the 321-node
// readout grid with a not-a-knot cubic spline, and integrating every bin
// exactly against exp(-i w t) (shower/api.py:461-467, readout.py:18-40).
// Spline and bin integral are both linear in the spectrum, so the host folds
// them into one complex basis M = K S (bins x table frequencies) once per
// grid, and one launch computes
//
//     bins[n, b] = Re sum_k M[b, k] * spectrum[k, n] * exp(-i w_k origin_n)
//
// as a real GEMM [Re M | -Im M] (bins x 2W) @ [Re; Im] (2W x observations) on
// MFMA (v_mfma_f32_16x16x4_f32). The demodulation is fused into the B-operand
// load; its phase w_k * origin_n is formed and reduced modulo 2 pi in FP64
// because origins reach thousands of ns, where an FP32 argument would lose
// ~1e-4 rad.
#include "lh_common.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 32;  // one wave computes 32 bins x 32 observations

__device__ __forceinline__ void demodulate(float2 s, double omega, double origin,
                                           float& re, float& im) {
    const double two_pi = 6.283185307179586476925286766559;
    double x = omega * origin;
    double reduced = x - two_pi * rint(x * (1.0 / two_pi));
    float sn, cs;
    sincosf(static_cast<float>(reduced), &sn, &cs);
    // s * exp(-i x)
    re = s.x * cs + s.y * sn;
    im = s.y * cs - s.x * sn;
}

__global__ __launch_bounds__(64) void readout_kernel(
    const float* __restrict__ basis,      // (n_bins, 2*n_freq): [Re M | -Im M]
    const double* __restrict__ omega,     // (n_freq)
    const float2* __restrict__ spectrum,  // (n_freq, n_obs)
    const double* __restrict__ origins,   // (n_obs) or null: no demodulation
    int n_bins, int n_freq, int n_obs,
    float* __restrict__ bins)             // (n_obs, n_bins)
{
    const int lane = threadIdx.x;
    const int r = lane & 15;
    const int kq = lane >> 4;
    const int row0 = blockIdx.x * kTile;
    const int col0 = blockIdx.y * kTile;
    const int stride = 2 * n_freq;

    int col[2] = {col0 + r, col0 + 16 + r};
    double origin[2];
#pragma unroll
    for (int jn = 0; jn < 2; ++jn)
        origin[jn] = (origins != nullptr && col[jn] < n_obs) ? origins[col[jn]] : 0.0;

    floatx4 acc[2][2];
#pragma unroll
    for (int im = 0; im < 2; ++im)
#pragma unroll
        for (int jn = 0; jn < 2; ++jn) acc[im][jn] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < n_freq; k0 += 4) {
        const int k = k0 + kq;
        const bool kin = k < n_freq;
        float a_re[2], a_im[2], b_re[2], b_im[2];
#pragma unroll
        for (int im = 0; im < 2; ++im) {
            const int row = row0 + 16 * im + r;
            const bool ok = kin && row < n_bins;
            a_re[im] = ok ? basis[(size_t)row * stride + k] : 0.f;
            a_im[im] = ok ? basis[(size_t)row * stride + n_freq + k] : 0.f;
        }
        const double w = kin ? omega[k] : 0.0;
#pragma unroll
        for (int jn = 0; jn < 2; ++jn) {
            b_re[jn] = 0.f;
            b_im[jn] = 0.f;
            if (kin && col[jn] < n_obs) {
                const float2 s = spectrum[(size_t)k * n_obs + col[jn]];
                if (origins != nullptr) {
                    demodulate(s, w, origin[jn], b_re[jn], b_im[jn]);
                } else {
                    b_re[jn] = s.x;
                    b_im[jn] = s.y;
                }
            }
        }
#pragma unroll
        for (int im = 0; im < 2; ++im)
#pragma unroll
            for (int jn = 0; jn < 2; ++jn) {
                acc[im][jn] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_re[im], b_re[jn],
                                                                   acc[im][jn], 0, 0, 0);
                acc[im][jn] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_im[im], b_im[jn],
                                                                   acc[im][jn], 0, 0, 0);
            }
    }

#pragma unroll
    for (int im = 0; im < 2; ++im)
#pragma unroll
        for (int jn = 0; jn < 2; ++jn)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = row0 + 16 * im + 4 * kq + i;
                const int c = col0 + 16 * jn + r;
                if (row < n_bins && c < n_obs) bins[(size_t)c * n_bins + row] = acc[im][jn][i];
            }
}

// The same product in split precision: basis and demodulated spectrum as FP16
// hi + lo pairs (x = hi + lo, lo the FP16 rounding of x - hi), every row of
// the basis and every observation scaled by a power of two so its largest
// value lies in [2^13, 2^14), and the three products hi*hi + hi*lo + lo*hi
// accumulated in FP32 on v_mfma_f32_16x16x16_f16 (the lo*lo term is below FP32
// rounding). The phase is formed and reduced in FP64 as above.
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_half(float x, _Float16& hi, _Float16& lo) {
    hi = static_cast<_Float16>(x);
    lo = static_cast<_Float16>(x - static_cast<float>(hi));
}

__global__ __launch_bounds__(64) void readout_split_kernel(
    const _Float16* __restrict__ basis_hi,   // (n_bins, 2*n_freq), rows scaled by row_scale
    const _Float16* __restrict__ basis_lo,
    const float* __restrict__ row_scale,     // (n_bins) powers of two
    const double* __restrict__ omega, const float2* __restrict__ spectrum,
    const double* __restrict__ origins, int n_bins, int n_freq, int n_obs,
    float* __restrict__ bins)
{
    const int lane = threadIdx.x;
    const int r = lane & 15;
    const int kq = lane >> 4;
    const int row0 = blockIdx.x * kTile;
    const int col0 = blockIdx.y * kTile;
    const int stride = 2 * n_freq;

    int col[2] = {col0 + r, col0 + 16 + r};
    double origin[2];
    float col_scale[2];
#pragma unroll
    for (int jn = 0; jn < 2; ++jn) {
        const bool ok = col[jn] < n_obs;
        origin[jn] = (origins != nullptr && ok) ? origins[col[jn]] : 0.0;
        // the observation's largest |re|, |im| (demodulation keeps |s|)
        float m = 0.f;
        if (ok)
            for (int k = kq; k < n_freq; k += 4) {
                const float2 s = spectrum[(size_t)k * n_obs + col[jn]];
                m = fmaxf(m, fmaxf(fabsf(s.x), fabsf(s.y)));
            }
        m = fmaxf(m, __shfl_xor(m, 16, 64));
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        col_scale[jn] = m > 0.f ? ldexpf(1.f, 13 - ilogbf(m)) : 1.f;
    }

    floatx4 acc[2][2];
#pragma unroll
    for (int im = 0; im < 2; ++im)
#pragma unroll
        for (int jn = 0; jn < 2; ++jn) acc[im][jn] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < n_freq; k0 += 16) {
        halfx4 a_re_hi[2], a_re_lo[2], a_im_hi[2], a_im_lo[2];
#pragma unroll
        for (int im = 0; im < 2; ++im) {
            const int row = row0 + 16 * im + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = k0 + 4 * kq + i;
                const bool ok = k < n_freq && row < n_bins;
                const size_t at = (size_t)row * stride + k;
                a_re_hi[im][i] = ok ? basis_hi[at] : (_Float16)0.f;
                a_re_lo[im][i] = ok ? basis_lo[at] : (_Float16)0.f;
                a_im_hi[im][i] = ok ? basis_hi[at + n_freq] : (_Float16)0.f;
                a_im_lo[im][i] = ok ? basis_lo[at + n_freq] : (_Float16)0.f;
            }
        }
        halfx4 b_re_hi[2], b_re_lo[2], b_im_hi[2], b_im_lo[2];
#pragma unroll
        for (int jn = 0; jn < 2; ++jn) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = k0 + 4 * kq + i;
                float re = 0.f, im = 0.f;
                if (k < n_freq && col[jn] < n_obs) {
                    const float2 s = spectrum[(size_t)k * n_obs + col[jn]];
                    if (origins != nullptr) {
                        demodulate(s, omega[k], origin[jn], re, im);
                    } else {
                        re = s.x;
                        im = s.y;
                    }
                }
                _Float16 h, l;
                split_half(re * col_scale[jn], h, l);
                b_re_hi[jn][i] = h;
                b_re_lo[jn][i] = l;
                split_half(im * col_scale[jn], h, l);
                b_im_hi[jn][i] = h;
                b_im_lo[jn][i] = l;
            }
        }
#pragma unroll
        for (int im = 0; im < 2; ++im)
#pragma unroll
            for (int jn = 0; jn < 2; ++jn) {
                floatx4 c = acc[im][jn];
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_re_hi[im], b_re_hi[jn], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_re_hi[im], b_re_lo[jn], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_re_lo[im], b_re_hi[jn], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_im_hi[im], b_im_hi[jn], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_im_hi[im], b_im_lo[jn], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16f16(a_im_lo[im], b_im_hi[jn], c, 0, 0, 0);
                acc[im][jn] = c;
            }
    }

#pragma unroll
    for (int im = 0; im < 2; ++im)
#pragma unroll
        for (int jn = 0; jn < 2; ++jn)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = row0 + 16 * im + 4 * kq + i;
                const int c = col0 + 16 * jn + r;
                if (row < n_bins && c < n_obs)
                    bins[(size_t)c * n_bins + row] =
                        acc[im][jn][i] / (row_scale[row] * col_scale[jn]);
            }
}

}  // namespace

extern "C" int lh_readout_bins_split(const void* basis_hi, const void* basis_lo,
                                     const float* row_scale, const double* omega,
                                     const void* spectrum, const double* origins, int n_bins,
                                     int n_freq, int n_obs, float* bins, void* stream) {
    LH_REQUIRE(n_bins > 0 && n_freq > 0 && n_obs >= 0,
               "lh_readout_bins_split: need n_bins > 0, n_freq > 0, n_obs >= 0 (got %d, %d, %d)",
               n_bins, n_freq, n_obs);
    if (n_obs == 0) return LH_OK;
    LH_REQUIRE(basis_hi && basis_lo && row_scale && omega && spectrum && bins,
               "lh_readout_bins_split: null device pointer");
    dim3 grid((n_bins + kTile - 1) / kTile, (n_obs + kTile - 1) / kTile);
    lh::KernelMark mark("readout_split_kernel", static_cast<hipStream_t>(stream));
    hipLaunchKernelGGL(readout_split_kernel, grid, dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<const _Float16*>(basis_hi),
                       static_cast<const _Float16*>(basis_lo), row_scale, omega,
                       static_cast<const float2*>(spectrum), origins, n_bins, n_freq, n_obs,
                       bins);
    return lh::check_launch("lh_readout_bins_split");
}

extern "C" int lh_readout_bins(const float* basis, const double* omega, const void* spectrum,
                               const double* origins, int n_bins, int n_freq, int n_obs,
                               float* bins, void* stream) {
    LH_REQUIRE(n_bins > 0 && n_freq > 0 && n_obs >= 0,
               "lh_readout_bins: need n_bins > 0, n_freq > 0, n_obs >= 0 (got %d, %d, %d)",
               n_bins, n_freq, n_obs);
    if (n_obs == 0) return LH_OK;
    LH_REQUIRE(basis && omega && spectrum && bins, "lh_readout_bins: null device pointer");
    dim3 grid((n_bins + kTile - 1) / kTile, (n_obs + kTile - 1) / kTile);
    lh::KernelMark mark("readout_kernel", static_cast<hipStream_t>(stream));
    hipLaunchKernelGGL(readout_kernel, grid, dim3(64), 0, static_cast<hipStream_t>(stream),
                       basis, omega, static_cast<const float2*>(spectrum), origins, n_bins,
                       n_freq, n_obs, bins);
    return lh::check_launch("lh_readout_bins");
}
