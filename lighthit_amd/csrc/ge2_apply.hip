// Stage (b): order >=2 apply of the band-folded RTE kernel pack.
//
// Equation (*) of the reference (directional.py:17-37) as fused in
// _apply_multi_spectral (_shower_basis.py:895-1050): for every OM d and
// source point p inside the radial window,
//
//   out[w, d] += sum_band e^{i w ref_band} scale(d,p,band) e^{i w r/v_band}
//                sum_l sum_a rot_l[a](p, band, w) cs_l[a](d,p, band, w)
//   rot_l[a] = sum_b E_l[a, b](d,p) ch[band, p, l-block + b, w]
//   cs_l[a]  = sum_pairs (alpha Y)(d,p) cubic(coef[band, j(d,p), l, pair, :, w], t(d,p))
//
// Two launch shapes share the geometry code:
//   * wide (many frequencies): one workgroup per OM. Its threads first build,
//     in FP64, the per-(OM, point) record (Wigner rotation rows E_l of every
//     degree, detector harmonics times acceptance, spline interval and
//     offset, attenuation) of a chunk of points into LDS; then each thread
//     owns one (band, frequency) lane and sums the chunk's points, reading
//     the records as LDS broadcasts and the pack and channels with coalesced
//     loads along the frequency axis. Bands are combined in a fixed order at
//     the end, so the result does not depend on scheduling.
//   * narrow (one frequency, the zero-frequency charge): one thread per
//     (OM, point) advances the Wigner recurrences degree by degree in
//     registers; the workgroup reduces its points in a fixed tree order.
// The heavy contraction is FP32; geometry, spline offset and the flight
// phase w r / v (up to ~10^3 rad) are formed in FP64 and reduced modulo 2 pi.
#include "lh_common.h"

#include <algorithm>

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBands = 9;
constexpr int kGroup = 8;     // points per degree-outer group in the wide kernel
constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Ge2Params {
    const float2* coef;        // (bands, intervals, degree+1, pairs, 4, cstride) complex64, or
                               // the bfloat16 hi + lo MFMA layout (coefficient_format 1)
    const unsigned* coef16;    // unused
    const double* log_r;       // (intervals+1)
    const double* absorption;  // (bands)
    const double* speed;       // (bands)
    const double* omega;       // (nfreq..)
    const double* alpha;       // (A+1)
    const double* points;      // (npoints, 3) frame coordinates
    const float2* channels;    // (bands, npoints, nchan, chstride)
    const double* band_ref;    // (bands) emission-time reference of each band [ns]
    const double* receivers;   // (nrec, 3) frame coordinates
    const double* looks;       // (nrec, 3) frame coordinates
    const double* out_scale;   // (bands, nrec)
    float2* out;               // (nfreq, nrec)
    int bands, intervals, degree, cstride, flight;
    int npoints, nchan, src_az, chstride;
    int nrec, nfreq;
    double r_low, r_high;
    int chunk, record_words;
    int nsplit, span;          // wide: points split over nsplit workgroups of `span` points
    float2* partial;           // wide with nsplit > 1: (nrec, nsplit, bands * nfreq)
};

__device__ __forceinline__ double parity(int n) { return (n & 1) ? -1.0 : 1.0; }

// Pack coefficient i (complex64; F is kept for a second storage form).
template <int F>
__device__ __forceinline__ float2 coef_at(const Ge2Params& P, size_t i) {
    static_assert(F == 0, "complex64 coefficients only");
    return P.coef[i];
}

__device__ __forceinline__ double ipow(double x, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

// sqrt(binom(2j, j - nu)), exact in double for the small degrees used here.
__device__ __forceinline__ double start_norm(int j, int nu) {
    double b = 1.0;
    const int k = j - nu;
    for (int i = 1; i <= k; ++i) b = b * (2 * j - k + i) / i;
    return sqrt(b);
}

// d^j_(mu nu) at j = max(|mu|, |nu|) (_directional_numba.py:36-49).
__device__ double start_value(int j, int mu, int nu, double half_cos, double half_sin) {
    double sign = 1.0;
    if (abs(nu) == j && abs(mu) != j) {
        const int tmp = mu;
        mu = nu;
        nu = tmp;
        sign *= parity(mu - nu);
    }
    if (mu == -j) {
        mu = j;
        nu = -nu;
        sign *= parity(j + nu);
    }
    return sign * start_norm(j, nu) * ipow(half_cos, j + nu) * ipow(-half_sin, j - nu);
}

// Real orthonormal harmonics of a vector up to degree A, flat l*l+l+m
// (_directional_numba.py:91-126).
template <int A>
__device__ void harmonics(double x, double y, double z, double (&out)[(A + 1) * (A + 1)]) {
    const double radius = sqrt(x * x + y * y + z * z);
    const double mu = z / radius;
    const double transverse = sqrt(x * x + y * y);
    const double sine = transverse / radius;
    const double cp = transverse == 0.0 ? 1.0 : x / transverse;
    const double sp = transverse == 0.0 ? 0.0 : y / transverse;
    double cm = 1.0, sm = 0.0;
    double diagonal = 0.28209479177387814347403972578039;  // 1/sqrt(4 pi)
#pragma unroll
    for (int m = 0; m <= A; ++m) {
        if (m) {
            diagonal *= sqrt((2.0 * m + 1) / (2.0 * m)) * sine;
            const double ncm = cm * cp - sm * sp;
            const double nsm = sm * cp + cm * sp;
            cm = ncm;
            sm = nsm;
        }
        double v0 = diagonal, v1 = 0.0;
#pragma unroll
        for (int ell = m; ell <= A; ++ell) {
            double value;
            if (ell == m) {
                value = v0;
            } else if (ell == m + 1) {
                value = sqrt(2.0 * m + 3) * mu * v0;
                v1 = value;
            } else {
                const double aa = sqrt((4.0 * ell * ell - 1) / ((double)ell * ell - m * m));
                const double bb = sqrt(((2.0 * ell + 1) * ((double)(ell - 1) * (ell - 1) - m * m))
                                       / ((2.0 * ell - 3) * ((double)ell * ell - m * m)));
                value = aa * mu * v1 - bb * v0;
                v0 = v1;
                v1 = value;
            }
            const int centre = ell * ell + ell;
            if (m == 0) {
                out[centre] = value;
            } else {
                out[centre - m] = 1.4142135623730950488 * value * sm;
                out[centre + m] = 1.4142135623730950488 * value * cm;
            }
        }
    }
}

// Wigner d^l_(mu nu)(theta) for mu = n in [0, M], nu = +-p with p in [0, A],
// advanced one degree at a time (the l-recurrence of _wigner_into,
// _directional_numba.py:52-88), and the real rotation rows E_l built from them
// (_shower_basis.py:1006-1029).
template <int A, int M>
struct Rotation {
    double cosine, half_cos, half_sin;
    double cosn[M + 1], sinn[M + 1];
    double cur[A + 1][M + 1][2], prev[A + 1][M + 1][2];

    __device__ void init(double c, double phi) {
        cosine = c;
        half_cos = sqrt((1.0 + c) / 2.0);
        half_sin = sqrt((1.0 - c) / 2.0);
#pragma unroll
        for (int n = 0; n <= M; ++n) {
            cosn[n] = cos(n * phi);
            sinn[n] = sin(n * phi);
        }
#pragma unroll
        for (int p = 0; p <= A; ++p)
#pragma unroll
            for (int n = 0; n <= M; ++n)
#pragma unroll
                for (int s = 0; s < 2; ++s) cur[p][n][s] = prev[p][n][s] = 0.0;
    }

    // Bring every recurrence to degree l (call with l = 0, 1, 2, ...).
    __device__ void step(int l) {
#pragma unroll
        for (int p = 0; p <= A; ++p)
#pragma unroll
            for (int n = 0; n <= M; ++n)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (p == 0 && s == 1) continue;
                    const int nu = s ? -p : p;
                    const int mu = n;
                    const int start = p > n ? p : n;
                    if (l < start) {
                        cur[p][n][s] = 0.0;
                    } else if (l == start) {
                        cur[p][n][s] = start_value(start, mu, nu, half_cos, half_sin);
                        prev[p][n][s] = 0.0;
                    } else {
                        const int ell = l - 1;
                        double value;
                        if (ell == 0) {
                            value = cosine * cur[p][n][s];
                        } else {
                            double lower = (double)(ell * ell - mu * mu) * (double)(ell * ell - nu * nu);
                            lower = lower > 0.0 ? sqrt(lower) : 0.0;
                            const double upper = sqrt((double)((ell + 1) * (ell + 1) - mu * mu)
                                                      * (double)((ell + 1) * (ell + 1) - nu * nu));
                            value = ((2 * ell + 1) * (ell * (ell + 1) * cosine - mu * nu) * cur[p][n][s]
                                     - (ell + 1) * lower * prev[p][n][s]) / (ell * upper);
                        }
                        prev[p][n][s] = cur[p][n][s];
                        cur[p][n][s] = value;
                    }
                }
    }

    // E_l[a][b], a = A + nu, b = M + m (after step(l)).
    __device__ void rows(int l, float (&E)[2 * A + 1][2 * M + 1]) const {
        const double root2 = 1.4142135623730950488;
#pragma unroll
        for (int a = 0; a < 2 * A + 1; ++a)
#pragma unroll
            for (int b = 0; b < 2 * M + 1; ++b) E[a][b] = 0.f;
        const int top_n = l < M ? l : M;
        const int top_p = l < A ? l : A;
        E[A][M] = (float)cur[0][0][0];
#pragma unroll
        for (int n = 1; n <= M; ++n) {
            if (n > top_n) continue;
            const double base = root2 * parity(n) * cur[0][n][0];
            E[A][M + n] = (float)(base * cosn[n]);
            E[A][M - n] = (float)(base * sinn[n]);
        }
#pragma unroll
        for (int p = 1; p <= A; ++p) {
            if (p > top_p) continue;
            E[A + p][M] = (float)(root2 * parity(p) * cur[p][0][0]);
#pragma unroll
            for (int n = 1; n <= M; ++n) {
                if (n > top_n) continue;
                const double cross = parity(p) * cur[p][n][0];
                const double straight = cur[p][n][1];
                const double cc = parity(n) * (cross + straight);
                const double ss = parity(n) * (cross - straight);
                E[A + p][M + n] = (float)(cc * cosn[n]);
                E[A + p][M - n] = (float)(cc * sinn[n]);
                E[A - p][M + n] = (float)(-ss * sinn[n]);
                E[A - p][M - n] = (float)(ss * cosn[n]);
            }
        }
    }
};

// Geometry of one (OM, point) pair in FP64. Returns false outside the window.
template <int A, int M>
struct PairGeometry {
    double radius, t;
    int j;
    double cosine, phi;
    float pos[(A + 1) * (A + 2) / 2], neg[(A + 1) * (A + 2) / 2];

    __device__ bool compute(const Ge2Params& P, int d, int p) {
        const double dx = P.receivers[3 * d + 0] - P.points[3 * p + 0];
        const double dy = P.receivers[3 * d + 1] - P.points[3 * p + 1];
        const double dz = P.receivers[3 * d + 2] - P.points[3 * p + 2];
        radius = sqrt(dx * dx + dy * dy + dz * dz);
        if (radius < P.r_low || radius > P.r_high) return false;
        cosine = dz / radius;
        cosine = cosine > 1.0 ? 1.0 : (cosine < -1.0 ? -1.0 : cosine);
        phi = atan2(dy, dx);
        const double transverse = sqrt(dx * dx + dy * dy);
        const double sine = transverse / radius;
        const double cp = transverse == 0.0 ? 1.0 : dx / transverse;
        const double sp = transverse == 0.0 ? 0.0 : dy / transverse;
        const double lx = P.looks[3 * d + 0], ly = P.looks[3 * d + 1], lz = P.looks[3 * d + 2];
        const double a1 = cosine * cp * lx + cosine * sp * ly - sine * lz;
        const double a2 = -sp * lx + cp * ly;
        const double a3 = sine * cp * lx + sine * sp * ly + cosine * lz;
        double Y[(A + 1) * (A + 1)];
        harmonics<A>(a1, a2, a3, Y);
        int k = 0;
#pragma unroll
        for (int lam = 0; lam <= A; ++lam)
#pragma unroll
            for (int mu = 0; mu <= lam; ++mu, ++k) {
                const int c = lam * lam + lam;
                pos[k] = (float)(P.alpha[lam] * Y[c + mu]);
                neg[k] = mu ? (float)(P.alpha[lam] * Y[c - mu]) : 0.f;
            }
        const double logr = log(radius);
        // searchsorted(log_r, logr, 'right') - 1, clamped to the intervals
        int lo = 0, hi = P.intervals + 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (P.log_r[mid] <= logr) lo = mid + 1; else hi = mid;
        }
        j = lo - 1;
        j = j < 0 ? 0 : (j > P.intervals - 1 ? P.intervals - 1 : j);
        t = logr - P.log_r[j];
        return true;
    }
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 phase_of(double angle) {
    const double reduced = angle - kTwoPi * rint(angle * (1.0 / kTwoPi));
    float s, c;
    sincosf((float)reduced, &s, &c);
    return make_float2(c, s);
}

// Record layout (32-bit words) of one (OM, point) in LDS for the wide kernel.
template <int A, int M>
struct Record {
    static constexpr int NA = 2 * A + 1, NM = 2 * M + 1, NPAIR = (A + 1) * (A + 2) / 2;
    static __host__ __device__ int words(int degree, int bands) {
        // valid, j, t, (pad), radius (2 words), G, scale, E; padded to 4 words
        const int w = 6 + 2 * NPAIR + bands + (degree + 1) * NA * NM;
        return (w + 3) & ~3;
    }
};

template <int A, int M, int F>
__global__ __launch_bounds__(kThreads, 2) void ge2_wide(Ge2Params P) {
    using R = Record<A, M>;
    constexpr int NA = R::NA, NM = R::NM, NPAIR = R::NPAIR;
    extern __shared__ float lds[];
    const int d = blockIdx.x;
    const int split = blockIdx.y;
    const int tid = threadIdx.x;
    const int lanes = P.bands * P.nfreq;
    // Per-(band, frequency) accumulators, each owned by one thread, then the
    // chunk's point records.
    float2* acc = reinterpret_cast<float2*>(lds);
    float* smem = lds + ((2 * lanes + 3) & ~3);
    for (int lw = tid; lw < lanes; lw += kThreads) acc[lw] = make_float2(0.f, 0.f);
    const int first = split * P.span;
    const int last = min(P.npoints, first + P.span);

    for (int base = first; base < last; base += P.chunk) {
        const int count = min(P.chunk, last - base);
        __syncthreads();
        for (int i = tid; i < count; i += kThreads) {
            float* rec = smem + (size_t)i * P.record_words;
            int* irec = reinterpret_cast<int*>(rec);
            PairGeometry<A, M> g;
            if (!g.compute(P, d, base + i)) {
                irec[0] = 0;
                continue;
            }
            irec[0] = 1;
            irec[1] = g.j;
            rec[2] = (float)g.t;
            reinterpret_cast<double*>(rec)[2] = g.radius;  // words 4-5... see below
#pragma unroll
            for (int k = 0; k < NPAIR; ++k) {
                rec[6 + 2 * k] = g.pos[k];
                rec[7 + 2 * k] = g.neg[k];
            }
            float* scale = rec + 6 + 2 * NPAIR;
            for (int lam = 0; lam < P.bands; ++lam)
                scale[lam] = (float)(P.out_scale[(size_t)lam * P.nrec + d]
                                     * exp(-P.absorption[lam] * g.radius)
                                     / (4.0 * 3.14159265358979323846 * g.radius * g.radius));
            float* E = scale + P.bands;
            Rotation<A, M> rot;
            rot.init(g.cosine, g.phi);
            for (int l = 0; l <= P.degree; ++l) {
                rot.step(l);
                float El[NA][NM];
                rot.rows(l, El);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int b = 0; b < NM; ++b) E[(l * NA + a) * NM + b] = El[a][b];
            }
        }
        __syncthreads();

        for (int lw = tid; lw < lanes; lw += kThreads) {
            const int lam = lw / P.nfreq;
            const int w = lw - lam * P.nfreq;
            const double inv_speed_w = P.flight ? P.omega[w] / P.speed[lam] : 0.0;
            const size_t cl = (size_t)NPAIR * 4 * P.cstride;  // per degree
            float2 total = acc[lw];
            // Points in groups of kGroup with degree outermost: the pack
            // coefficients of one degree are loaded once per run of points
            // in the same radial interval and reused across the group.
            for (int g0 = 0; g0 < count; g0 += kGroup) {
                float2 accp[kGroup];
#pragma unroll
                for (int q = 0; q < kGroup; ++q) accp[q] = make_float2(0.f, 0.f);
                int starts = 0;
                for (int l = 0; l <= P.degree; ++l) {
                    const int top_p = l < A ? l : A;
                    const int keep = l < M ? l : M;
                    const int centre = starts + keep;
                    starts += 2 * (l < P.src_az ? l : P.src_az) + 1;
                    int jc = -1;
                    float2 cf[NPAIR][4];
#pragma unroll
                    for (int q = 0; q < kGroup; ++q) {
                        const int i = g0 + q;
                        if (i >= count) break;
                        const float* rec = smem + (size_t)i * P.record_words;
                        const int* irec = reinterpret_cast<const int*>(rec);
                        if (irec[0] == 0) continue;
                        const int j = irec[1];
                        if (j != jc) {
                            const size_t cdeg
                                = (((size_t)lam * P.intervals + j) * (P.degree + 1) + l) * cl + w;
#pragma unroll
                            for (int k = 0; k < NPAIR; ++k)
#pragma unroll
                                for (int pw = 0; pw < 4; ++pw)
                                    cf[k][pw] = coef_at<F>(P, cdeg + (size_t)(k * 4 + pw)
                                                                     * P.cstride);
                            jc = j;
                        }
                        const float t = rec[2];
                        const float* G = rec + 6;
                        const float* El = rec + 6 + 2 * NPAIR + P.bands + l * NA * NM;
                        const float2* chbase = P.channels
                            + ((size_t)lam * P.npoints + base + i) * P.nchan * P.chstride + w;
                        float2 chv[NM];
#pragma unroll
                        for (int bb = 0; bb < NM; ++bb) {
                            const int m = bb - M;
                            chv[bb] = (m >= -keep && m <= keep)
                                          ? chbase[(size_t)(centre + m) * P.chstride]
                                          : make_float2(0.f, 0.f);
                        }
                        float2 cs[NA];
#pragma unroll
                        for (int aa = 0; aa < NA; ++aa) cs[aa] = make_float2(0.f, 0.f);
                        int k = 0;
#pragma unroll
                        for (int dl = 0; dl <= A; ++dl)
#pragma unroll
                            for (int mu = 0; mu <= dl; ++mu, ++k) {
                                if (mu > top_p) continue;
                                float2 radial;
                                radial.x = ((cf[k][0].x * t + cf[k][1].x) * t + cf[k][2].x) * t + cf[k][3].x;
                                radial.y = ((cf[k][0].y * t + cf[k][1].y) * t + cf[k][2].y) * t + cf[k][3].y;
                                const float pos = G[2 * k], neg = G[2 * k + 1];
                                cs[A + mu].x += pos * radial.x;
                                cs[A + mu].y += pos * radial.y;
                                if (mu) {
                                    cs[A - mu].x += neg * radial.x;
                                    cs[A - mu].y += neg * radial.y;
                                }
                            }
#pragma unroll
                        for (int aa = 0; aa < NA; ++aa) {
                            const int nu = aa - A;
                            if (nu < -top_p || nu > top_p) continue;
                            float2 r = make_float2(0.f, 0.f);
#pragma unroll
                            for (int bb = 0; bb < NM; ++bb) {
                                const float e = El[aa * NM + bb];
                                r.x += e * chv[bb].x;
                                r.y += e * chv[bb].y;
                            }
                            const float2 v = cmul(r, cs[aa]);
                            accp[q].x += v.x;
                            accp[q].y += v.y;
                        }
                    }
                }
                // attenuation and flight phase of every point of the group
#pragma unroll
                for (int q = 0; q < kGroup; ++q) {
                    const int i = g0 + q;
                    if (i >= count) break;
                    const float* rec = smem + (size_t)i * P.record_words;
                    if (reinterpret_cast<const int*>(rec)[0] == 0) continue;
                    const double radius = reinterpret_cast<const double*>(rec)[2];
                    const float scale = rec[6 + 2 * NPAIR + lam];
                    float2 sp = make_float2(scale, 0.f);
                    if (P.flight) {
                        const float2 ph = phase_of(inv_speed_w * radius);
                        sp = make_float2(scale * ph.x, scale * ph.y);
                    }
                    const float2 v = cmul(accp[q], sp);
                    total.x += v.x;
                    total.y += v.y;
                }
            }
            acc[lw] = total;
        }
    }

    if (P.nsplit > 1) {
        // partial sums of this split; ge2_wide_reduce combines them in order
        __syncthreads();
        float2* out = P.partial + ((size_t)d * P.nsplit + split) * lanes;
        for (int lw = tid; lw < lanes; lw += kThreads) out[lw] = acc[lw];
        return;
    }
    // Combine the bands of every frequency in a fixed order.
    __syncthreads();
    float2* red = acc;
    for (int lw = tid; lw < lanes; lw += kThreads) {
        const int lam = lw / P.nfreq;
        const int w = lw - lam * P.nfreq;
        red[lw] = cmul(acc[lw], phase_of(P.omega[w] * P.band_ref[lam]));
    }
    __syncthreads();
    for (int w = tid; w < P.nfreq; w += kThreads) {
        float2 sum = make_float2(0.f, 0.f);
        for (int lam = 0; lam < P.bands; ++lam) {
            sum.x += red[lam * P.nfreq + w].x;
            sum.y += red[lam * P.nfreq + w].y;
        }
        P.out[(size_t)w * P.nrec + d] = sum;
    }
}

// Zero spectrum (no source points).
__global__ void ge2_zero_out(Ge2Params P) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < P.nrec * P.nfreq) P.out[idx] = make_float2(0.f, 0.f);
}

// out[w, d] = sum_band e^{i w ref_band} sum_split partial[d, split, band, w]
__global__ void ge2_wide_reduce(Ge2Params P) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P.nrec * P.nfreq) return;
    const int d = idx / P.nfreq, w = idx - d * P.nfreq;
    const int lanes = P.bands * P.nfreq;
    float2 sum = make_float2(0.f, 0.f);
    for (int lam = 0; lam < P.bands; ++lam) {
        float2 band = make_float2(0.f, 0.f);
        for (int c = 0; c < P.nsplit; ++c) {
            const float2 v = P.partial[((size_t)d * P.nsplit + c) * lanes + lam * P.nfreq + w];
            band.x += v.x;
            band.y += v.y;
        }
        const float2 v = cmul(band, phase_of(P.omega[w] * P.band_ref[lam]));
        sum.x += v.x;
        sum.y += v.y;
    }
    P.out[(size_t)w * P.nrec + d] = sum;
}

// Workgroups of the wide kernel per OM: enough to fill the chip, at least 16
// points each.
int wide_splits(int nrec, int npoints) {
    const int want = (1024 + nrec - 1) / (nrec > 0 ? nrec : 1);
    const int most = (npoints + 15) / 16;
    return std::max(1, std::min(want, most));
}

// One frequency: one thread per (OM, point); fixed-order tree reduction.
template <int A, int M, int F>
__global__ __launch_bounds__(kThreads) void ge2_narrow(Ge2Params P, int w) {
    constexpr int NA = 2 * A + 1, NM = 2 * M + 1;
    __shared__ float2 red[kMaxBands][kThreads];
    const int d = blockIdx.x;
    const int tid = threadIdx.x;
    float2 acc[kMaxBands];
#pragma unroll
    for (int b = 0; b < kMaxBands; ++b) acc[b] = make_float2(0.f, 0.f);

    for (int p = tid; p < P.npoints; p += kThreads) {
        PairGeometry<A, M> g;
        if (!g.compute(P, d, p)) continue;
        const float t = (float)g.t;
        Rotation<A, M> rot;
        rot.init(g.cosine, g.phi);
        float2 resp[kMaxBands];
#pragma unroll
        for (int b = 0; b < kMaxBands; ++b) resp[b] = make_float2(0.f, 0.f);
        const size_t cl = (size_t)((A + 1) * (A + 2) / 2) * 4 * P.cstride;
        int starts = 0;
        for (int l = 0; l <= P.degree; ++l) {
            rot.step(l);
            float El[NA][NM];
            rot.rows(l, El);
            const int top_p = l < A ? l : A;
            const int keep = l < M ? l : M;
            const int centre = starts + keep;
            starts += 2 * (l < P.src_az ? l : P.src_az) + 1;
#pragma unroll
            for (int lam = 0; lam < kMaxBands; ++lam) {
                if (lam >= P.bands) break;
                const float2* chbase = P.channels + ((size_t)lam * P.npoints + p) * P.nchan * P.chstride + w;
                float2 chv[NM];
#pragma unroll
                for (int b = 0; b < NM; ++b) {
                    const int m = b - M;
                    chv[b] = (m >= -keep && m <= keep) ? chbase[(size_t)(centre + m) * P.chstride]
                                                       : make_float2(0.f, 0.f);
                }
                float2 cs[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) cs[a] = make_float2(0.f, 0.f);
                const size_t cdeg = ((size_t)lam * P.intervals + g.j) * (P.degree + 1) * cl
                                    + (size_t)l * cl + w;
                int k = 0;
#pragma unroll
                for (int dl = 0; dl <= A; ++dl)
#pragma unroll
                    for (int mu = 0; mu <= dl; ++mu, ++k) {
                        if (mu > top_p) continue;
                        const size_t c = cdeg + (size_t)k * 4 * P.cstride;
                        const float2 c0 = coef_at<F>(P, c), c1 = coef_at<F>(P, c + P.cstride),
                                     c2 = coef_at<F>(P, c + 2 * P.cstride),
                                     c3 = coef_at<F>(P, c + 3 * P.cstride);
                        float2 radial;
                        radial.x = ((c0.x * t + c1.x) * t + c2.x) * t + c3.x;
                        radial.y = ((c0.y * t + c1.y) * t + c2.y) * t + c3.y;
                        cs[A + mu].x += g.pos[k] * radial.x;
                        cs[A + mu].y += g.pos[k] * radial.y;
                        if (mu) {
                            cs[A - mu].x += g.neg[k] * radial.x;
                            cs[A - mu].y += g.neg[k] * radial.y;
                        }
                    }
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int nu = a - A;
                    if (nu < -top_p || nu > top_p) continue;
                    float2 r = make_float2(0.f, 0.f);
#pragma unroll
                    for (int b = 0; b < NM; ++b) {
                        r.x += El[a][b] * chv[b].x;
                        r.y += El[a][b] * chv[b].y;
                    }
                    const float2 v = cmul(r, cs[a]);
                    resp[lam].x += v.x;
                    resp[lam].y += v.y;
                }
            }
        }
#pragma unroll
        for (int lam = 0; lam < kMaxBands; ++lam) {
            if (lam >= P.bands) break;
            const double scale = P.out_scale[(size_t)lam * P.nrec + d] * exp(-P.absorption[lam] * g.radius)
                                 / (4.0 * 3.14159265358979323846 * g.radius * g.radius);
            float2 sp = make_float2((float)scale, 0.f);
            if (P.flight) {
                const float2 ph = phase_of(P.omega[w] * g.radius / P.speed[lam]);
                sp = make_float2((float)scale * ph.x, (float)scale * ph.y);
            }
            const float2 v = cmul(resp[lam], sp);
            acc[lam].x += v.x;
            acc[lam].y += v.y;
        }
    }
#pragma unroll
    for (int lam = 0; lam < kMaxBands; ++lam) red[lam][tid] = acc[lam];
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (tid < half) {
#pragma unroll
            for (int lam = 0; lam < kMaxBands; ++lam) {
                red[lam][tid].x += red[lam][tid + half].x;
                red[lam][tid].y += red[lam][tid + half].y;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        float2 sum = make_float2(0.f, 0.f);
        for (int lam = 0; lam < P.bands; ++lam) {
            const float2 v = cmul(red[lam][0], phase_of(P.omega[w] * P.band_ref[lam]));
            sum.x += v.x;
            sum.y += v.y;
        }
        P.out[(size_t)w * P.nrec + d] = sum;
    }
}

constexpr int kLdsBudget = 64 * 1024;

template <int A, int M, int F>
int launch(Ge2Params P, hipStream_t stream) {
    if (P.nfreq == 1) {
        {
            lh::KernelMark mark("ge2_narrow", stream);
            hipLaunchKernelGGL((ge2_narrow<A, M, F>), dim3(P.nrec), dim3(kThreads), 0, stream, P,
                               0);
        }
        return lh::check_launch("lh_ge2_apply (narrow)");
    }
    P.record_words = Record<A, M>::words(P.degree, P.bands);
    const int reduce_words = (2 * P.bands * P.nfreq + 3) & ~3;
    LH_REQUIRE(reduce_words * 4 <= kLdsBudget / 2, "lh_ge2_apply: bands x frequencies too large");
    P.chunk = std::min(64, (kLdsBudget - reduce_words * 4) / (4 * P.record_words));
    LH_REQUIRE(P.chunk >= 1, "lh_ge2_apply: degree too large for the LDS record");
    const size_t lds = (size_t)(reduce_words + P.chunk * P.record_words) * 4;
    P.nsplit = wide_splits(P.nrec, P.npoints);
    P.span = (P.npoints + P.nsplit - 1) / P.nsplit;
    {
        lh::KernelMark mark("ge2_wide", stream);
        hipLaunchKernelGGL((ge2_wide<A, M, F>), dim3(P.nrec, P.nsplit), dim3(kThreads), lds, stream,
                           P);
    }
    int status = lh::check_launch("lh_ge2_apply (wide)");
    if (status != LH_OK || P.nsplit == 1) return status;
    const int total = P.nrec * P.nfreq;
    hipLaunchKernelGGL(ge2_wide_reduce, dim3((total + 255) / 256), dim3(256), 0, stream, P);
    return lh::check_launch("lh_ge2_apply (wide reduce)");
}

// ------------------------------------------------------------------ MFMA form
//
// The precision form of the multi-frequency apply. For one (OM d, point p) pair
// in radial interval j and one degree l, the rotation of the point's channels
// and the detector-side pair sums fold into one real matrix Q_l:
//   sum_a rot_l[a] cs_l[a] = sum_b ch_l[b] H_l[b],   H_l[b] = sum_K Q_l[b, K] C_jl[K],
//   Q_l[b, (k, pw)] = t^(3 - pw) (E_l[A + mu_k][b] pos_k + E_l[A - mu_k][b] neg_k),
// with C_jl[K] (K = pair k x cubic power pw) the interval's spline coefficients
// of every (band, frequency) column: a product whose right operand is shared
// by every pair in interval j. Pairs are grouped by interval; a 16-row MFMA
// tile holds 16 / R pairs, R the power of two >= 2M + 1 (rows b = -M..M of
// each, then zero rows), so that every pair's rows fall on the same lanes of
// the result whatever its place in the tile and its sums over b run in one
// order. Columns are (band, frequency). Q and C are split into bfloat16 hi + lo and
// hi*hi + hi*lo + lo*hi is accumulated in FP32 on v_mfma_f32_16x16x16_bf16
// (the lo*lo term lies below the FP32 sum's rounding). The epilogue multiplies
// by the point's channels, sums b and l in a fixed order, applies attenuation
// and the FP64-reduced flight phase and writes the pair's spectrum; the last
// kernel sums every OM's points in point order, so the result depends on
// neither the grouping nor the scheduling.

typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaChunk = 4;   // column tiles of 16 per pass

// Device workspace of the MFMA form (laid out by mfma_layout).
struct MfmaWork {
    int* count;       // (intervals) pairs per interval
    int* offset;      // (intervals + 1) first pair slot of each interval
    int* cursor;      // (intervals) unused
    int* count_jd;    // (intervals, npoints) pairs per (interval, point)
    int* offset_jd;   // (intervals, npoints) first slot of each (interval, point)
    int* tile_off;    // (intervals + 1) first tile of each interval
    int* totals;      // [pairs, tiles]
    unsigned char* valid;   // (nrec, npoints) pair inside the radial window
    int* list;        // (pairs) d * npoints + p, grouped by interval
    float* records;   // (pairs, record_words)
    float2* partial;  // (nrec, npoints, bands * nfreq)
};

__host__ __device__ inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Sizes (and, with base != nullptr, pointers) of the MFMA workspace.
__host__ size_t mfma_layout(const Ge2Params& P, int words, char* base, MfmaWork* W) {
    const size_t pairs = (size_t)P.nrec * P.npoints;
    const int J = P.intervals;
    size_t at = 0;
    auto take = [&](size_t bytes) { const size_t here = at; at += align256(bytes); return here; };
    const size_t o_count = take(sizeof(int) * J), o_offset = take(sizeof(int) * (J + 1));
    const size_t o_cursor = take(sizeof(int) * J), o_tile = take(sizeof(int) * (J + 1));
    const size_t o_count_jd = take(sizeof(int) * (size_t)J * P.npoints);
    const size_t o_offset_jd = take(sizeof(int) * (size_t)J * P.npoints);
    const size_t o_totals = take(sizeof(int) * 2), o_valid = take(pairs);
    const size_t o_list = take(sizeof(int) * pairs);
    const size_t o_records = take(sizeof(float) * pairs * words);
    const size_t o_partial = take(sizeof(float2) * pairs * P.bands * P.nfreq);
    if (base != nullptr) {
        W->count = reinterpret_cast<int*>(base + o_count);
        W->offset = reinterpret_cast<int*>(base + o_offset);
        W->cursor = reinterpret_cast<int*>(base + o_cursor);
        W->count_jd = reinterpret_cast<int*>(base + o_count_jd);
        W->offset_jd = reinterpret_cast<int*>(base + o_offset_jd);
        W->tile_off = reinterpret_cast<int*>(base + o_tile);
        W->totals = reinterpret_cast<int*>(base + o_totals);
        W->valid = reinterpret_cast<unsigned char*>(base + o_valid);
        W->list = reinterpret_cast<int*>(base + o_list);
        W->records = reinterpret_cast<float*>(base + o_records);
        W->partial = reinterpret_cast<float2*>(base + o_partial);
    }
    return at;
}

// Radial interval of one (OM, point) pair; false outside the window.
__device__ bool pair_interval(const Ge2Params& P, int d, int p, int& j) {
    const double dx = P.receivers[3 * d + 0] - P.points[3 * p + 0];
    const double dy = P.receivers[3 * d + 1] - P.points[3 * p + 1];
    const double dz = P.receivers[3 * d + 2] - P.points[3 * p + 2];
    const double radius = sqrt(dx * dx + dy * dy + dz * dz);
    if (radius < P.r_low || radius > P.r_high) return false;
    const double logr = log(radius);
    int lo = 0, hi = P.intervals + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (P.log_r[mid] <= logr) lo = mid + 1; else hi = mid;
    }
    j = lo - 1;
    j = j < 0 ? 0 : (j > P.intervals - 1 ? P.intervals - 1 : j);
    return true;
}

// One workgroup per source point: the window flag of each OM and the point's
// pairs per radial interval (LDS histogram), stored interval-major.
__global__ void ge2_mfma_count(Ge2Params P, MfmaWork W) {
    extern __shared__ int hist[];
    const int p = blockIdx.x;
    for (int j = threadIdx.x; j < P.intervals; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    for (int d = threadIdx.x; d < P.nrec; d += blockDim.x) {
        int j = 0;
        const bool inside = pair_interval(P, d, p, j);
        W.valid[(size_t)d * P.npoints + p] = inside ? 1 : 0;
        if (inside) atomicAdd(&hist[j], 1);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < P.intervals; j += blockDim.x)
        W.count_jd[(size_t)j * P.npoints + p] = hist[j];
}

// Exclusive scan of the (interval, point) counts in interval-major order (one
// workgroup of 1024 threads, each over a contiguous run), then the pair and
// unit offsets of every interval.
__global__ __launch_bounds__(1024) void ge2_mfma_scan(Ge2Params P, MfmaWork W, int per_unit) {
    __shared__ int sums[1024];
    const int n = P.intervals * P.npoints;
    const int t = threadIdx.x;
    const int run = (n + 1023) / 1024;
    const int lo = min(n, t * run), hi = min(n, lo + run);
    int local = 0;
    for (int i = lo; i < hi; ++i) local += W.count_jd[i];
    sums[t] = local;
    __syncthreads();
    for (int step = 1; step < 1024; step <<= 1) {   // inclusive Hillis-Steele scan
        const int add = t >= step ? sums[t - step] : 0;
        __syncthreads();
        sums[t] += add;
        __syncthreads();
    }
    int running = sums[t] - local;
    for (int i = lo; i < hi; ++i) {
        W.offset_jd[i] = running;
        running += W.count_jd[i];
    }
    __syncthreads();
    if (t == 0) {
        int units = 0;
        for (int j = 0; j < P.intervals; ++j) {
            const int begin = W.offset_jd[(size_t)j * P.npoints];
            const int end = j + 1 < P.intervals ? W.offset_jd[(size_t)(j + 1) * P.npoints]
                                                : sums[1023];
            W.offset[j] = begin;
            W.count[j] = end - begin;
            W.tile_off[j] = units;
            units += (end - begin + per_unit - 1) / per_unit;
        }
        W.offset[P.intervals] = sums[1023];
        W.tile_off[P.intervals] = units;
        W.totals[0] = sums[1023];
        W.totals[1] = units;
    }
}

// One workgroup per source point: its pairs into the list, grouped by
// interval, points in order inside an interval, so that neighbouring pairs
// share the point's channels. The order of one point's OMs inside an interval
// follows LDS atomics; it changes which pairs share a tile, never a pair's
// result.
__global__ void ge2_mfma_fill(Ge2Params P, MfmaWork W) {
    extern __shared__ int cursor[];
    const int p = blockIdx.x;
    for (int j = threadIdx.x; j < P.intervals; j += blockDim.x)
        cursor[j] = W.offset_jd[(size_t)j * P.npoints + p];
    __syncthreads();
    for (int d = threadIdx.x; d < P.nrec; d += blockDim.x) {
        const size_t pair = (size_t)d * P.npoints + p;
        if (!W.valid[pair]) continue;
        int j = 0;
        pair_interval(P, d, p, j);
        W.list[atomicAdd(&cursor[j], 1)] = (int)pair;
    }
}

// FP64 geometry of every listed pair into its record (the wide kernel's LDS
// record, in global memory): d * npoints + p, interval, offset t, radius,
// detector pair factors, band scales and the rotation rows of every degree.
template <int A, int M>
__global__ void ge2_mfma_records(Ge2Params P, MfmaWork W) {
    constexpr int NA = 2 * A + 1, NM = 2 * M + 1, NPAIR = (A + 1) * (A + 2) / 2;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= W.totals[0]) return;
    const int idx = W.list[slot];
    const int d = idx / P.npoints, p = idx - d * P.npoints;
    PairGeometry<A, M> g;
    g.compute(P, d, p);
    float* rec = W.records + (size_t)slot * P.record_words;
    int* irec = reinterpret_cast<int*>(rec);
    irec[0] = idx;
    irec[1] = g.j;
    rec[2] = (float)g.t;
    reinterpret_cast<double*>(rec)[2] = g.radius;
#pragma unroll
    for (int k = 0; k < NPAIR; ++k) {
        rec[6 + 2 * k] = g.pos[k];
        rec[7 + 2 * k] = g.neg[k];
    }
    float* scale = rec + 6 + 2 * NPAIR;
    for (int lam = 0; lam < P.bands; ++lam)
        scale[lam] = (float)(P.out_scale[(size_t)lam * P.nrec + d] * exp(-P.absorption[lam] * g.radius)
                             / (4.0 * 3.14159265358979323846 * g.radius * g.radius));
    float* E = scale + P.bands;
    Rotation<A, M> rot;
    rot.init(g.cosine, g.phi);
    for (int l = 0; l <= P.degree; ++l) {
        rot.step(l);
        float El[NA][NM];
        rot.rows(l, El);
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NM; ++b) E[(l * NA + a) * NM + b] = El[a][b];
    }
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// q = hi + lo, both bfloat16 rounded to nearest even (v_cvt_pk_bf16_f32); the
// remainder of the pair is below 2^-16 of q.
__device__ __forceinline__ void split_bf16(floatx4 q, shortx4& hi, shortx4& lo) {
    const bf16x4 h = __builtin_convertvector(q, bf16x4);
    const floatx4 rest = q - __builtin_convertvector(h, floatx4);
    hi = __builtin_bit_cast(shortx4, h);
    lo = __builtin_bit_cast(shortx4, __builtin_convertvector(rest, bf16x4));
}

// Canonical pair k -> (lambda, mu): lambda = 0..A, mu = 0..lambda.
template <int A>
__device__ __forceinline__ int pair_mu(int k) {
    int lam = 0;
    while ((lam + 1) * (lam + 2) / 2 <= k) ++lam;
    return k - lam * (lam + 1) / 2;
}

// Rows per pair (a power of two >= 2M + 1), pairs per 16-row tile, tiles per
// wave unit (they share every load of the right operand).
template <int M> struct MfmaShape {
    static constexpr int NM = 2 * M + 1;
    static constexpr int RPP = NM == 1 ? 1 : (NM <= 4 ? 4 : (NM <= 8 ? 8 : 16));
    static constexpr int PPT = 16 / RPP;
    static constexpr int MT = 2;
    static constexpr int UNIT = PPT * MT;
};

template <int A, int M>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ge2_mfma_apply(Ge2Params P, MfmaWork W,
                                                       const shortx4* __restrict__ B) {
    constexpr int NA = 2 * A + 1, NM = 2 * M + 1, NPAIR = (A + 1) * (A + 2) / 2;
    constexpr int KQ = (NPAIR + 3) / 4 * 4, KSTEPS = KQ / 4;
    using S = MfmaShape<M>;
    constexpr int RPP = S::RPP, PPT = S::PPT, MT = S::MT, UNIT = S::UNIT;
    constexpr int SLOTS = RPP == 1 ? 4 : 1;      // pairs one lane's four result rows hold
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int group = lane >> 4;                 // K quarter (A, B) / row quarter (D)
    const int column = lane & 15;
    const int ncol = P.bands * P.nfreq;
    const int ntile_col = (ncol + 15) / 16;
    const int nchunk = (ntile_col + kMfmaChunk - 1) / kMfmaChunk;
    const int total_items = W.totals[1] * nchunk;
    const int J = P.intervals;
    const int arow_q = column / RPP, arow_b = column - arow_q * RPP;   // operand A row
    const int first_q = (4 * group) / RPP;                             // result rows
    const size_t point_stride = (size_t)P.nchan * P.chstride;
    // operand A of this lane: pair k = 4 s + group of row arow_b, read from
    // clamped addresses and masked
    const int arow_bc = arow_b < NM ? arow_b : NM - 1;
    int kpos[KSTEPS], epos[KSTEPS], eneg[KSTEPS], kmu[KSTEPS];
    bool kok[KSTEPS];
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) {
        const int k = 4 * s + group;
        kok[s] = k < NPAIR;
        const int kc = kok[s] ? k : 0;
        const int mu = pair_mu<A>(kc);
        kmu[s] = mu;
        kpos[s] = 6 + 2 * kc;
        epos[s] = (A + mu) * NM + arow_bc;
        eneg[s] = (A - mu) * NM + arow_bc;
    }
    const int e_first = 6 + 2 * NPAIR + P.bands;

    // A work item is one unit of pairs and one chunk of column tiles; the
    // chunks of a unit go to neighbouring waves, which share its records and
    // channels in cache. Workgroups are dealt to the 8 XCDs in turn (grid a
    // multiple of 8): each XCD takes a contiguous eighth of the items, so the
    // units of one radial interval, which read the same coefficients, meet in
    // one L2.
    const int xcd = blockIdx.x % 8, rank = blockIdx.x / 8, per_xcd = gridDim.x / 8;
    const int first_item = (int)((long long)total_items * xcd / 8);
    const int end_item = (int)((long long)total_items * (xcd + 1) / 8);
    for (int item = first_item + rank * 4 + wave; item < end_item; item += per_xcd * 4) {
        const int unit = item / nchunk;
        const int c0 = (item - unit * nchunk) * kMfmaChunk;
        const int nct = min(kMfmaChunk, ntile_col - c0);
        // interval of the unit: the last j with tile_off[j] <= unit
        int lo = 0, hi = J;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (W.tile_off[mid] <= unit) lo = mid; else hi = mid;
        }
        const int j = lo;
        const int first = W.offset[j] + (unit - W.tile_off[j]) * UNIT;
        const int npairs = min(UNIT, W.offset[j] + W.count[j] - first);
        // operand A rows (one pair per tile and lane) and, for each result row
        // of this lane, the offset of its point's channels (clamped to a valid
        // address; rows without a pair or beyond 2M + 1 are masked)
        const float* arec[MT];
        bool arow_ok[MT];
        floatx4 tpow[MT];
        size_t chrow[MT][4];
        bool row_ok[MT][4];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int qa = mt * PPT + arow_q;
            arow_ok[mt] = qa < npairs && arow_b < NM;
            arec[mt] = W.records + (size_t)(first + (qa < npairs ? qa : 0)) * P.record_words;
            const float t = arec[mt][2];
            tpow[mt] = floatx4{t * t * t, t * t, t, 1.f};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 4 * group + i;
                const int q = mt * PPT + row / RPP;
                const int b = row % RPP;
                const bool ok = q < npairs && b < NM;
                const int pt = ok ? W.list[first + q] % P.npoints : 0;
                row_ok[mt][i] = ok;
                chrow[mt][i] = (size_t)pt * point_stride;
            }
        }

        float2 tot[MT][kMfmaChunk][SLOTS];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int c = 0; c < kMfmaChunk; ++c)
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) tot[mt][c][s] = make_float2(0.f, 0.f);
        int starts = 0;
        for (int l = 0; l <= P.degree; ++l) {
            const int top_p = l < A ? l : A;
            const int keep = l < M ? l : M;
            const int centre = starts + keep;
            starts += 2 * (l < P.src_az ? l : P.src_az) + 1;
            // Q_l fragments (row arow of every tile; K = 4 s + group, power i)
            const bool row_deg = arow_b - M >= -keep && arow_b - M <= keep;
            shortx4 qhi[MT][KSTEPS], qlo[MT][KSTEPS];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const float* E = arec[mt] + e_first + l * NA * NM;
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s) {
                    const float pos = E[epos[s]] * arec[mt][kpos[s]];
                    const float neg = E[eneg[s]] * arec[mt][kpos[s] + 1];
                    const bool ok = arow_ok[mt] && kok[s] && row_deg && kmu[s] <= top_p;
                    const float base = ok ? (kmu[s] ? pos + neg : pos) : 0.f;
                    split_bf16(base * tpow[mt], qhi[mt][s], qlo[mt][s]);
                }
            }
            // rows of this degree: |b - M| <= min(l, M); every row reads a
            // channel of this degree's block (clamped), masked when unused
            bool use[MT][4];
            size_t chcol[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int bb = (4 * group + i) % RPP - M;
                const int inside = bb < -keep ? -keep : (bb > keep ? keep : bb);
                chcol[i] = (size_t)(centre + inside) * P.chstride;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    use[mt][i] = row_ok[mt][i] && bb >= -keep && bb <= keep;
            }
#pragma unroll
            for (int c = 0; c < kMfmaChunk; ++c) {
                if (c >= nct) break;
                const int col = 16 * (c0 + c) + column;
                const bool col_ok = col < ncol;
                const int lam = col_ok ? col / P.nfreq : 0;
                const int w = col_ok ? col - lam * P.nfreq : 0;
                // channels of every result row, issued before the products
                const float2* chl = P.channels + (size_t)lam * P.npoints * point_stride + w;
                float2 ch[MT][4];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) ch[mt][i] = chl[chrow[mt][i] + chcol[i]];
                // operand B: loaded once, used by every tile of the unit
                shortx4 b_re_hi[KSTEPS], b_re_lo[KSTEPS], b_im_hi[KSTEPS], b_im_lo[KSTEPS];
                const shortx4* bl = B + ((((((size_t)lam * J + j) * (P.degree + 1) + l) * P.cstride
                                          + w) * KQ) + group) * 4;
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s) {
                    const shortx4* src = bl + 16 * s;
                    b_re_hi[s] = src[0];
                    b_re_lo[s] = src[1];
                    b_im_hi[s] = src[2];
                    b_im_lo[s] = src[3];
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    if (mt * PPT >= npairs) break;
                    floatx4 hre = {0.f, 0.f, 0.f, 0.f}, him = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int s = 0; s < KSTEPS; ++s) {
                        hre = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qhi[mt][s], b_re_hi[s], hre, 0, 0, 0);
                        him = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qhi[mt][s], b_im_hi[s], him, 0, 0, 0);
                        hre = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qhi[mt][s], b_re_lo[s], hre, 0, 0, 0);
                        him = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qhi[mt][s], b_im_lo[s], him, 0, 0, 0);
                        hre = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qlo[mt][s], b_re_hi[s], hre, 0, 0, 0);
                        him = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(qlo[mt][s], b_im_hi[s], him, 0, 0, 0);
                    }
                    // epilogue: H of this lane's rows times the point's channels
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int s = RPP == 1 ? i : 0;
                        const float2 v = use[mt][i] ? ch[mt][i] : make_float2(0.f, 0.f);
                        tot[mt][c][s].x += v.x * hre[i] - v.y * him[i];
                        tot[mt][c][s].y += v.x * him[i] + v.y * hre[i];
                    }
                }
            }
        }
        // sum the four row quarters of every pair, then scale and phase
#pragma unroll
        for (int c = 0; c < kMfmaChunk; ++c) {
            if (c >= nct) break;
            const int col = 16 * (c0 + c) + column;
            const bool col_ok = col < ncol;
            const int lam = col_ok ? col / P.nfreq : 0;
            const int w = col_ok ? col - lam * P.nfreq : 0;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                if (mt * PPT >= npairs) break;
#pragma unroll
                for (int q = 0; q < PPT; ++q) {
                    float2 v = make_float2(0.f, 0.f);
#pragma unroll
                    for (int s = 0; s < SLOTS; ++s)
                        if (q - first_q == s) v = tot[mt][c][s];
                    v.x += __shfl_xor(v.x, 16, 64);
                    v.y += __shfl_xor(v.y, 16, 64);
                    v.x += __shfl_xor(v.x, 32, 64);
                    v.y += __shfl_xor(v.y, 32, 64);
                    const int qq = mt * PPT + q;
                    if (group != 0 || !col_ok || qq >= npairs) continue;
                    const float* rec = W.records + (size_t)(first + qq) * P.record_words;
                    const int idx = reinterpret_cast<const int*>(rec)[0];
                    const double radius = reinterpret_cast<const double*>(rec)[2];
                    const float scale = rec[6 + 2 * NPAIR + lam];
                    float2 sp = make_float2(scale, 0.f);
                    if (P.flight) {
                        const float2 ph = phase_of(P.omega[w] / P.speed[lam] * radius);
                        sp = make_float2(scale * ph.x, scale * ph.y);
                    }
                    W.partial[(size_t)idx * ncol + col] = cmul(v, sp);
                }
            }
        }
    }
}

// out[w, d] = sum_band e^{i w ref_band} sum_p partial[d, p, band, w], points in
// order: one workgroup per OM, a thread per (band, frequency) column, then the
// bands of every frequency in order.
__global__ void ge2_mfma_reduce(Ge2Params P, MfmaWork W) {
    extern __shared__ float2 band_sum[];
    const int d = blockIdx.x;
    const int ncol = P.bands * P.nfreq;
    const unsigned char* valid = W.valid + (size_t)d * P.npoints;
    const float2* part = W.partial + (size_t)d * P.npoints * ncol;
    for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
        float2 sum = make_float2(0.f, 0.f);
        int p = 0;
        for (; p + 4 <= P.npoints; p += 4) {
            float2 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                v[k] = valid[p + k] ? part[(size_t)(p + k) * ncol + col] : make_float2(0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sum.x += v[k].x;
                sum.y += v[k].y;
            }
        }
        for (; p < P.npoints; ++p) {
            if (!valid[p]) continue;
            const float2 v = part[(size_t)p * ncol + col];
            sum.x += v.x;
            sum.y += v.y;
        }
        band_sum[col] = sum;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < P.nfreq; w += blockDim.x) {
        float2 sum = make_float2(0.f, 0.f);
        for (int lam = 0; lam < P.bands; ++lam) {
            const float2 v = cmul(band_sum[lam * P.nfreq + w], phase_of(P.omega[w] * P.band_ref[lam]));
            sum.x += v.x;
            sum.y += v.y;
        }
        P.out[(size_t)w * P.nrec + d] = sum;
    }
}

template <int A, int M>
int launch_mfma(Ge2Params P, const void* workspace, size_t workspace_bytes, hipStream_t stream) {
    constexpr int NM = 2 * M + 1;
    P.record_words = Record<A, M>::words(P.degree, P.bands);
    MfmaWork W{};
    const size_t need = mfma_layout(P, P.record_words, nullptr, nullptr);
    LH_REQUIRE(workspace != nullptr && workspace_bytes >= need,
               "lh_ge2_apply: workspace of %zu bytes, %zu needed", workspace_bytes, need);
    mfma_layout(P, P.record_words, static_cast<char*>(const_cast<void*>(workspace)), &W);
    const size_t pairs = (size_t)P.nrec * P.npoints;
    if (pairs == 0) {
        const int total = P.nrec * P.nfreq;
        hipLaunchKernelGGL(ge2_zero_out, dim3((total + 255) / 256), dim3(256), 0, stream, P);
        return lh::check_launch("lh_ge2_apply (mfma zero)");
    }
    const unsigned grid_pairs = (unsigned)((pairs + 255) / 256);
    const size_t lds_j = sizeof(int) * (size_t)P.intervals;
    hipLaunchKernelGGL(ge2_mfma_count, dim3(P.npoints), dim3(256), lds_j, stream, P, W);
    hipLaunchKernelGGL(ge2_mfma_scan, dim3(1), dim3(1024), 0, stream, P, W, MfmaShape<M>::UNIT);
    hipLaunchKernelGGL(ge2_mfma_fill, dim3(P.npoints), dim3(256), lds_j, stream, P, W);
    hipLaunchKernelGGL((ge2_mfma_records<A, M>), dim3(grid_pairs), dim3(256), 0, stream, P, W);
    int status = lh::check_launch("lh_ge2_apply (mfma set-up)");
    if (status != LH_OK) return status;
    const size_t max_tiles = (pairs + MfmaShape<M>::UNIT - 1) / MfmaShape<M>::UNIT + P.intervals;
    const size_t chunks = ((size_t)P.bands * P.nfreq + 16 * kMfmaChunk - 1) / (16 * kMfmaChunk);
    // a multiple of 8 workgroups (the kernel deals items per XCD)
    const size_t want = std::min<size_t>((max_tiles * chunks + 3) / 4, 4096);
    const unsigned groups = (unsigned)std::max<size_t>(8, (want + 7) / 8 * 8);
    {
        lh::KernelMark mark("ge2_mfma_apply", stream);
        hipLaunchKernelGGL((ge2_mfma_apply<A, M>), dim3(groups), dim3(256), 0, stream, P, W,
                           reinterpret_cast<const shortx4*>(P.coef));
    }
    status = lh::check_launch("lh_ge2_apply (mfma)");
    if (status != LH_OK) return status;
    const int total = P.nrec * P.nfreq;
    hipLaunchKernelGGL(ge2_mfma_reduce, dim3(P.nrec), dim3(256),
                       sizeof(float2) * (size_t)P.bands * P.nfreq, stream, P, W);
    return lh::check_launch("lh_ge2_apply (mfma reduce)");
}

template <int A>
int dispatch_mfma(int m, const Ge2Params& P, const void* ws, size_t bytes, hipStream_t s) {
    switch (m) {
        case 0: return launch_mfma<A, 0>(P, ws, bytes, s);
        case 1: return launch_mfma<A, 1>(P, ws, bytes, s);
        case 2: return launch_mfma<A, 2>(P, ws, bytes, s);
        case 3: return launch_mfma<A, 3>(P, ws, bytes, s);
        case 4: return launch_mfma<A, 4>(P, ws, bytes, s);
        default: break;
    }
    lh::set_error("lh_ge2_apply: azimuthal degree %d not built (0..4)", m);
    return LH_ERR_ARGUMENT;
}

// Records of the MFMA form for acceptance degree A and azimuthal degree m.
int mfma_record_words(int A, int m, int degree, int bands) {
    switch (A * 8 + m) {
        case 16: return Record<2, 0>::words(degree, bands);
        case 17: return Record<2, 1>::words(degree, bands);
        case 18: return Record<2, 2>::words(degree, bands);
        case 19: return Record<2, 3>::words(degree, bands);
        case 20: return Record<2, 4>::words(degree, bands);
        case 24: return Record<3, 0>::words(degree, bands);
        case 25: return Record<3, 1>::words(degree, bands);
        case 26: return Record<3, 2>::words(degree, bands);
        case 27: return Record<3, 3>::words(degree, bands);
        case 28: return Record<3, 4>::words(degree, bands);
        default: return 0;
    }
}

template <int A, int F>
int dispatch_m(int m, const Ge2Params& P, hipStream_t s) {
    switch (m) {
        case 0: return launch<A, 0, F>(P, s);
        case 1: return launch<A, 1, F>(P, s);
        case 2: return launch<A, 2, F>(P, s);
        case 3: return launch<A, 3, F>(P, s);
        case 4: return launch<A, 4, F>(P, s);
        default: break;
    }
    lh::set_error("lh_ge2_apply: azimuthal degree %d not built (0..4)", m);
    return LH_ERR_ARGUMENT;
}

}  // namespace

extern "C" size_t lh_ge2_workspace_bytes(int bands, int freq, int receivers, int points) {
    if (freq <= 1 || receivers <= 0) return 0;
    const int nsplit = wide_splits(receivers, points);
    return nsplit > 1 ? sizeof(float2) * (size_t)receivers * nsplit * bands * freq : 0;
}

extern "C" size_t lh_ge2_apply_workspace_bytes(const lh_ge2_pack* pack,
                                               const lh_ge2_source* source, int receivers,
                                               int freq) {
    if (pack == nullptr || source == nullptr || receivers <= 0) return 0;
    if (pack->coefficient_format != 1)
        return lh_ge2_workspace_bytes(pack->bands, freq, receivers, source->points);
    Ge2Params P{};
    P.nrec = receivers;
    P.npoints = source->points;
    P.intervals = pack->intervals;
    P.bands = pack->bands;
    P.nfreq = freq;
    const int m = std::min(source->azimuthal_degree, pack->degree);
    const int words = mfma_record_words(pack->acceptance_degree, m, pack->degree, pack->bands);
    return words > 0 ? mfma_layout(P, words, nullptr, nullptr) : 0;
}

extern "C" int lh_ge2_apply(const lh_ge2_pack* pack, const lh_ge2_source* source,
                            const lh_ge2_receivers* receivers, const double* alpha, int nfreq,
                            double radial_low, double radial_high, void* out, void* workspace,
                            size_t workspace_bytes, void* stream) {
    LH_REQUIRE(pack && source && receivers, "lh_ge2_apply: null descriptor");
    LH_REQUIRE(receivers->count >= 0, "lh_ge2_apply: negative receiver count");
    if (receivers->count == 0) return LH_OK;
    LH_REQUIRE(pack->bands >= 1 && pack->bands <= kMaxBands, "lh_ge2_apply: 1..%d bands supported",
               kMaxBands);
    LH_REQUIRE(pack->intervals >= 1 && pack->degree >= 0, "lh_ge2_apply: empty pack");
    LH_REQUIRE(nfreq >= 1 && nfreq <= pack->freq_stride && nfreq <= source->freq_stride,
               "lh_ge2_apply: %d frequencies requested, pack holds %d and source %d", nfreq,
               pack->freq_stride, source->freq_stride);
    LH_REQUIRE(source->degree >= pack->degree, "source degree is insufficient");
    int expected = 0;
    for (int l = 0; l <= source->degree; ++l)
        expected += 2 * std::min(l, source->azimuthal_degree) + 1;
    LH_REQUIRE(source->channels == expected,
               "lh_ge2_apply: %d source channels, degree %d / azimuthal %d needs %d",
               source->channels, source->degree, source->azimuthal_degree, expected);
    LH_REQUIRE(source->points >= 0, "lh_ge2_apply: negative point count");
    LH_REQUIRE(radial_low < radial_high, "radial window must lie inside the kernel cache");
    LH_REQUIRE(pack->coefficients && pack->log_r && pack->absorption && pack->speed && pack->omega
                   && alpha && receivers->positions && receivers->looks && receivers->out_scale
                   && out && (source->points == 0 || (source->points_m && source->channels_data
                                                      && source->reference_ns)),
               "lh_ge2_apply: null device pointer");
    LH_REQUIRE(pack->coefficient_format == 0 || pack->coefficient_format == 1,
               "lh_ge2_apply: coefficient_format must be 0 (complex64) or 1 (bfloat16 hi + lo, "
               "MFMA layout)");
    Ge2Params P{};
    P.coef = static_cast<const float2*>(pack->coefficients);
    P.coef16 = nullptr;
    P.log_r = pack->log_r;
    P.absorption = pack->absorption;
    P.speed = pack->speed;
    P.omega = pack->omega;
    P.alpha = alpha;
    P.points = source->points_m;
    P.channels = static_cast<const float2*>(source->channels_data);
    P.band_ref = source->reference_ns;
    P.receivers = receivers->positions;
    P.looks = receivers->looks;
    P.out_scale = receivers->out_scale;
    P.out = static_cast<float2*>(out);
    P.bands = pack->bands;
    P.intervals = pack->intervals;
    P.degree = pack->degree;
    P.cstride = pack->freq_stride;
    P.flight = pack->flight;
    P.npoints = source->points;
    P.nchan = source->channels;
    P.src_az = source->azimuthal_degree;
    P.chstride = source->freq_stride;
    P.nrec = receivers->count;
    P.nfreq = nfreq;
    P.r_low = radial_low;
    P.r_high = radial_high;
    const int m = std::min(source->azimuthal_degree, pack->degree);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (pack->coefficient_format == 1) {
        switch (pack->acceptance_degree) {
            case 2: return dispatch_mfma<2>(m, P, workspace, workspace_bytes, s);
            case 3: return dispatch_mfma<3>(m, P, workspace, workspace_bytes, s);
            default: break;
        }
        lh::set_error("lh_ge2_apply: acceptance degree %d not built (2, 3)",
                      pack->acceptance_degree);
        return LH_ERR_ARGUMENT;
    }
    const size_t need = lh_ge2_workspace_bytes(P.bands, nfreq, P.nrec, P.npoints);
    LH_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr),
               "lh_ge2_apply: workspace of %zu bytes, %zu needed", workspace_bytes, need);
    P.partial = static_cast<float2*>(workspace);
    switch (pack->acceptance_degree) {
        case 2: return dispatch_m<2, 0>(m, P, s);
        case 3: return dispatch_m<3, 0>(m, P, s);
        default: break;
    }
    lh::set_error("lh_ge2_apply: acceptance degree %d not built (2, 3)", pack->acceptance_degree);
    return LH_ERR_ARGUMENT;
}

// ---------------------------------------------------------------- near field
//
// Clamped order >=2 of element portions closer to an OM than the tables'
// radial minimum (near_field.py:91-173 with the fused-looks cone evaluator
// _cone_many_looks, muon/track_ge2_primitive.py:612-660). A portion is an
// axially symmetric point source moving along its element's axis; the OM sits
// at (impact, 0, u) of the portion frame for every quadrature node u. Points
// inside the radial minimum take the amplitude at the minimum along the same
// direction and the flight phase of the true distance. One workgroup per
// portion; the first thread sets up each node (Wigner column, detector
// harmonics, spline offset, band scales) in LDS; every thread then owns one
// frequency and sums nodes, bands, degrees and pairs in a fixed order (FP64).

namespace {

constexpr int kNearThreads = 128;   // frequencies per portion
constexpr int kNearMaxDegree = 16;
constexpr int kNearMaxBands = 32;

struct NearParams {
    lh_ge2_pack pack;
    lh_near_portions por;
    const double* alpha;
    double clamp_low, radial_low, radial_high;
    double* out;     // (count, nfreq, 2)
    int nfreq;
};

template <int A>
__global__ void __launch_bounds__(kNearThreads) near_field(NearParams P) {
    constexpr int kPairs = (A + 1) * (A + 2) / 2;
    __shared__ double s_k[kNearMaxDegree + 1][kPairs];   // rotation x alpha x detector harmonic
    __shared__ double s_scale[kNearMaxBands];
    __shared__ double s_radius, s_t, s_weight, s_u;
    __shared__ int s_j, s_skip;
    const int q = blockIdx.x;
    const int f = threadIdx.x;
    const lh_ge2_pack& pk = P.pack;
    const lh_near_portions& por = P.por;
    const int degree = pk.degree, bands = pk.bands;
    const double omega = f < P.nfreq ? pk.omega[f] : 0.0;
    const double clock = por.clock[q];
    double acc_re = 0.0, acc_im = 0.0;
    const double* harm = por.harmonics + (size_t)q * bands * (degree + 1);
    for (int node = por.node_start[q]; node < por.node_start[q + 1]; ++node) {
        if (threadIdx.x == 0) {
            const double impact = por.impact[q];
            const double u = por.u[node];
            const double radius = hypot(impact, u);
            int skip = radius < P.clamp_low || radius > P.radial_high || radius == 0.0;
            if (!skip) {
                const double cosine = u / radius, sine = impact / radius;
                const double amplitude = fmax(radius, P.radial_low);
                const double la = log(amplitude);
                int lo = 0, hi = pk.intervals;   // searchsorted(log_r, la, 'right') - 1
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (pk.log_r[mid] <= la) lo = mid; else hi = mid;
                }
                if (la < pk.log_r[0]) lo = 0;
                s_j = lo;
                s_t = la - pk.log_r[lo];
                s_radius = radius;
                for (int b = 0; b < bands; ++b)
                    s_scale[b] = por.weight[(size_t)q * bands + b]
                                 * exp(-pk.absorption[b] * amplitude)
                                 / (4.0 * M_PI * amplitude * amplitude);
                const double* look = por.look + 3 * (size_t)q;
                double y[(A + 1) * (A + 1)];
                harmonics<A>(cosine * look[0] - sine * look[2], look[1],
                             sine * look[0] + cosine * look[2], y);
                Rotation<A, 0> rot;
                rot.init(cosine, 0.0);
                for (int l = 0; l <= degree; ++l) {
                    rot.step(l);
                    int pair = 0;
#pragma unroll
                    for (int lam = 0; lam <= A; ++lam) {
#pragma unroll
                        for (int mu = 0; mu <= lam; ++mu, ++pair) {
                            double value = 0.0;
                            if (mu <= l) {
                                const double rotation =
                                    mu == 0 ? rot.cur[0][0][0]
                                            : 1.4142135623730950488 * parity(mu) * rot.cur[mu][0][0];
                                value = rotation * P.alpha[lam] * y[lam * lam + lam + mu];
                            }
                            s_k[l][pair] = value;
                        }
                    }
                }
                s_weight = por.w[node];
                s_u = u;
            }
            s_skip = skip;
        }
        __syncthreads();
        if (!s_skip && f < P.nfreq) {
            const double t = s_t;
            const int j = s_j;
            double node_re = 0.0, node_im = 0.0;
            for (int b = 0; b < bands; ++b) {
                double band_re = 0.0, band_im = 0.0;
                const float2* c = static_cast<const float2*>(pk.coefficients)
                                  + (((size_t)b * pk.intervals + j) * (degree + 1)) * kPairs * 4
                                        * pk.freq_stride;
                for (int l = 0; l <= degree; ++l) {
                    const double h = harm[b * (degree + 1) + l];
                    if (h == 0.0) continue;
                    double l_re = 0.0, l_im = 0.0;
#pragma unroll
                    for (int pair = 0; pair < kPairs; ++pair) {
                        const float2* cc = c + ((size_t)l * kPairs + pair) * 4 * pk.freq_stride + f;
                        const double re = (((double)cc[0].x * t + cc[pk.freq_stride].x) * t
                                           + cc[2 * pk.freq_stride].x) * t
                                          + cc[3 * pk.freq_stride].x;
                        const double im = (((double)cc[0].y * t + cc[pk.freq_stride].y) * t
                                           + cc[2 * pk.freq_stride].y) * t
                                          + cc[3 * pk.freq_stride].y;
                        l_re += s_k[l][pair] * re;
                        l_im += s_k[l][pair] * im;
                    }
                    band_re += h * l_re;
                    band_im += h * l_im;
                }
                const double angle = pk.flight ? omega * s_radius / pk.speed[b] : 0.0;
                double sa, ca;
                sincos(angle, &sa, &ca);
                const double pr = s_scale[b] * ca, pi = s_scale[b] * sa;
                node_re += band_re * pr - band_im * pi;
                node_im += band_re * pi + band_im * pr;
            }
            // quadrature weight and the emission-time shift e^{-i u clock omega}
            double sd, cd;
            sincos(-s_u * clock * omega, &sd, &cd);
            const double w = s_weight;
            acc_re += w * (node_re * cd - node_im * sd);
            acc_im += w * (node_re * sd + node_im * cd);
        }
        __syncthreads();
    }
    if (f < P.nfreq) {
        double sc, cc;
        sincos(omega * por.phase_ns[q], &sc, &cc);
        const double density = por.density[q];
        double* o = P.out + ((size_t)q * P.nfreq + f) * 2;
        o[0] = density * (acc_re * cc - acc_im * sc);
        o[1] = density * (acc_re * sc + acc_im * cc);
    }
}

}  // namespace

extern "C" int lh_near_field(const lh_ge2_pack* pack, const lh_near_portions* portions,
                             const double* alpha, int nfreq, double clamp_low, double radial_low,
                             double radial_high, double* out, void* stream) {
    LH_REQUIRE(pack && portions, "lh_near_field: null descriptor");
    if (portions->count == 0) return LH_OK;
    LH_REQUIRE(portions->count > 0, "lh_near_field: negative portion count");
    LH_REQUIRE(nfreq >= 1 && nfreq <= kNearThreads && nfreq <= pack->freq_stride,
               "lh_near_field: 1..%d frequencies", kNearThreads);
    LH_REQUIRE(pack->degree >= 0 && pack->degree <= kNearMaxDegree && pack->bands >= 1
                   && pack->bands <= kNearMaxBands && pack->intervals >= 1,
               "lh_near_field: degree <= %d and 1..%d bands", kNearMaxDegree, kNearMaxBands);
    LH_REQUIRE(pack->acceptance_degree >= 0 && pack->acceptance_degree <= 4,
               "lh_near_field: acceptance degree must lie in 0..4");
    LH_REQUIRE(pack->coefficient_format == 0, "lh_near_field: complex64 coefficients only");
    LH_REQUIRE(radial_low < radial_high && clamp_low <= radial_low,
               "lh_near_field: need clamp_low <= radial_low < radial_high");
    const lh_near_portions& p = *portions;
    LH_REQUIRE(p.impact && p.look && p.weight && p.harmonics && p.phase_ns && p.clock && p.density
                   && p.node_start && p.u && p.w && alpha && out && pack->coefficients
                   && pack->log_r && pack->absorption && pack->speed && pack->omega,
               "lh_near_field: null device pointer");
    NearParams P{};
    P.pack = *pack;
    P.por = p;
    P.alpha = alpha;
    P.clamp_low = clamp_low;
    P.radial_low = radial_low;
    P.radial_high = radial_high;
    P.out = out;
    P.nfreq = nfreq;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (pack->acceptance_degree) {
        case 0: near_field<0><<<p.count, kNearThreads, 0, s>>>(P); break;
        case 1: near_field<1><<<p.count, kNearThreads, 0, s>>>(P); break;
        case 2: near_field<2><<<p.count, kNearThreads, 0, s>>>(P); break;
        case 3: near_field<3><<<p.count, kNearThreads, 0, s>>>(P); break;
        default: near_field<4><<<p.count, kNearThreads, 0, s>>>(P); break;
    }
    return lh::check_launch("lh_near_field");
}
