/*
 * Oracle of stage (b): order >=2 apply of band-folded RTE kernels, FP64.
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 *
 * Restates _apply_multi_spectral (_shower_basis.py:895-1050) with its helpers
 * _wigner_into / _start_value / _harmonics_into / _parity
 * (_directional_numba.py:31-126) and _start_norm_table (:389-398), for the
 * route's tolerance_per_point = 0 (every band active, _shower_basis.py:711-725).
 * Channels and coefficients are complex64 as in the reference (channel_dtype,
 * fold_bundle dtype); all arithmetic and accumulation is complex128.
 * Parallel over receivers like the reference's prange; every receiver row is
 * summed by one thread in point order.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef float complex cf32;
typedef double complex cf64;

static double parity(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

/* _start_value (_directional_numba.py:36-49). */
static double start_value(int j, int mu, int nu, double half_cos, double half_sin,
                          const double* start_norm, int jmax) {
    double sign = 1.0;
    if (abs(nu) == j && abs(mu) != j) {
        int tmp = mu; mu = nu; nu = tmp;
        sign *= parity(mu - nu);
    }
    if (mu == -j) {
        mu = j;
        nu = -nu;
        sign *= parity(j + nu);
    }
    return sign * start_norm[j * (2 * jmax + 1) + nu + jmax]
        * pow(half_cos, j + nu) * pow(-half_sin, j - nu);
}

/* _wigner_into (_directional_numba.py:52-88): out[l][nu+max_out][mu]. */
static void wigner_into(int degree, int max_out, int max_in, double cosine,
                        const double* start_norm, int jmax, double* out) {
    int width = 2 * max_out + 1;
    memset(out, 0, sizeof(double) * (degree + 1) * width * (max_in + 1));
    double half_cos = sqrt((1.0 + cosine) / 2.0);
    double half_sin = sqrt((1.0 - cosine) / 2.0);
    for (int nu = -max_out; nu <= max_out; ++nu) {
        for (int mu = 0; mu <= max_in; ++mu) {
            int start = abs(nu) > mu ? abs(nu) : mu;
            if (start > degree) continue;
            double previous = 0.0;
            double current = start_value(start, mu, nu, half_cos, half_sin, start_norm, jmax);
            out[(start * width + nu + max_out) * (max_in + 1) + mu] = current;
            for (int ell = start; ell < degree; ++ell) {
                double value;
                if (ell == 0) {
                    value = cosine * current;
                } else {
                    double lower = (double)(ell * ell - mu * mu) * (double)(ell * ell - nu * nu);
                    lower = lower > 0.0 ? sqrt(lower) : 0.0;
                    double upper = sqrt((double)((ell + 1) * (ell + 1) - mu * mu)
                                        * (double)((ell + 1) * (ell + 1) - nu * nu));
                    value = ((2 * ell + 1) * (ell * (ell + 1) * cosine - mu * nu) * current
                             - (ell + 1) * lower * previous) / (ell * upper);
                }
                previous = current;
                current = value;
                out[((ell + 1) * width + nu + max_out) * (max_in + 1) + mu] = current;
            }
        }
    }
}

/* _harmonics_into (_directional_numba.py:91-126): flat l*l+l+m. */
static void harmonics_into(double x, double y, double z, int degree, double* out) {
    double radius = sqrt(x * x + y * y + z * z);
    double mu = z / radius;
    double transverse = sqrt(x * x + y * y);
    double sine = transverse / radius;
    double cp = transverse == 0.0 ? 1.0 : x / transverse;
    double sp = transverse == 0.0 ? 0.0 : y / transverse;
    double cm = 1.0, sm = 0.0;
    double diagonal = 1.0 / sqrt(4 * M_PI);
    double v0 = 0.0, v1 = 0.0;
    for (int m = 0; m <= degree; ++m) {
        if (m) {
            diagonal *= sqrt((2.0 * m + 1) / (2.0 * m)) * sine;
            double ncm = cm * cp - sm * sp;
            double nsm = sm * cp + cm * sp;
            cm = ncm;
            sm = nsm;
        }
        v0 = diagonal;
        for (int ell = m; ell <= degree; ++ell) {
            double value;
            if (ell == m) {
                value = v0;
            } else if (ell == m + 1) {
                value = sqrt(2.0 * m + 3) * mu * v0;
                v1 = value;
            } else {
                double aa = sqrt((4.0 * ell * ell - 1) / ((double)ell * ell - m * m));
                double bb = sqrt(((2.0 * ell + 1) * ((double)(ell - 1) * (ell - 1) - m * m))
                                 / ((2.0 * ell - 3) * ((double)ell * ell - m * m)));
                value = aa * mu * v1 - bb * v0;
                v0 = v1;
                v1 = value;
            }
            int centre = ell * ell + ell;
            if (m == 0) {
                out[centre] = value;
            } else {
                out[centre - m] = sqrt(2.0) * value * sm;
                out[centre + m] = sqrt(2.0) * value * cm;
            }
        }
    }
}

/* _start_norm_table (_directional_numba.py:389-398). */
void oracle_start_norm_table(int jmax, double* table) {
    for (int j = 0; j <= jmax; ++j)
        for (int nu = -jmax; nu <= jmax; ++nu)
            table[j * (2 * jmax + 1) + nu + jmax] =
                (abs(nu) <= j) ? exp(0.5 * (lgamma(2.0 * j + 1) - lgamma(j + nu + 1.0)
                                            - lgamma(j - nu + 1.0)))
                               : 0.0;
}

void oracle_wigner(int degree, int max_out, int max_in, double cosine, double* out) {
    int jmax = max_out > max_in ? max_out : max_in;
    double* norm = (double*)malloc(sizeof(double) * (jmax + 1) * (2 * jmax + 1));
    oracle_start_norm_table(jmax, norm);
    wigner_into(degree, max_out, max_in, cosine, norm, jmax, out);
    free(norm);
}

void oracle_harmonics(double x, double y, double z, int degree, double* out) {
    harmonics_into(x, y, z, degree, out);
}

/* The erow block of one degree (_shower_basis.py:1006-1029), dense
 * (2*max_out+1) x (2*max_m+1). */
static void rotation_rows(int ell, int max_out, int max_m, const double* wigner,
                          const double* cosines, const double* sines, double* erow) {
    int width = 2 * max_out + 1, wm = 2 * max_m + 1;
    const double root2 = sqrt(2.0);
#define W(l, nu, mu) wigner[((l) * width + (nu)) * (max_m + 1) + (mu)]
#define E(a, b) erow[(a) * wm + (b)]
    memset(erow, 0, sizeof(double) * width * wm);
    E(max_out, max_m) = W(ell, max_out, 0);
    int top_n = ell < max_m ? ell : max_m;
    int top_p = ell < max_out ? ell : max_out;
    for (int n = 1; n <= top_n; ++n) {
        double base = root2 * parity(n) * W(ell, max_out, n);
        E(max_out, max_m + n) = base * cosines[n];
        E(max_out, max_m - n) = base * sines[n];
    }
    for (int p = 1; p <= top_p; ++p) {
        E(max_out + p, max_m) = root2 * parity(p) * W(ell, max_out + p, 0);
        for (int n = 1; n <= top_n; ++n) {
            double cross = parity(p) * W(ell, max_out + p, n);
            double straight = W(ell, max_out - p, n);
            double cc = parity(n) * (cross + straight);
            double ss = parity(n) * (cross - straight);
            E(max_out + p, max_m + n) = cc * cosines[n];
            E(max_out + p, max_m - n) = cc * sines[n];
            E(max_out - p, max_m + n) = -ss * sines[n];
            E(max_out - p, max_m - n) = ss * cosines[n];
        }
    }
#undef W
#undef E
}

/* Dense rotation rows of every degree (the arithmetic of
 * rotation_rows_explicit, directional.py:721-774): out[l][a][b]. */
void oracle_rotation_rows(int degree, int max_out, int max_m, double cosine, double phi,
                          double* out) {
    int width = 2 * max_out + 1, wm = 2 * max_m + 1;
    double* wig = (double*)malloc(sizeof(double) * (degree + 1) * width * (max_m + 1));
    oracle_wigner(degree, max_out, max_m, cosine, wig);
    double cosines[64], sines[64];
    for (int n = 0; n <= max_m; ++n) {
        cosines[n] = cos(n * phi);
        sines[n] = sin(n * phi);
    }
    for (int l = 0; l <= degree; ++l)
        rotation_rows(l, max_out, max_m, wig, cosines, sines, out + (size_t)l * width * wm);
    free(wig);
}

/*
 * _apply_multi_spectral with tolerance_per_point = 0.
 * points (P,3), receivers (N,3), looks (N,3) in frame coordinates;
 * channels (nlam, P, nchan, nfreq) complex64; out_scale (nlam, N);
 * starts (>= degree+1); coefficients (nlam, nint, degree+1, npair, 4, nfreq)
 * complex64; result (N, nfreq) complex128.
 */
void oracle_apply_multi_spectral(
    const double* points, int npoints, const double* receivers, const double* looks, int nrec,
    const cf32* channels, int nchan, const double* out_scale, const int64_t* starts,
    int degree, int max_m, int max_out, const double* alpha, const int64_t* pair_lambda,
    const int64_t* pair_mu, int npair, const double* log_r, int nint, const cf32* coefficients,
    const double* absorption, const double* velocity, const double* omega, int nfreq,
    int nlam, int flight, double radial_low, double radial_high, cf64* result) {
    int jmax = max_m > max_out ? max_m : max_out;
    double* start_norm = (double*)malloc(sizeof(double) * (jmax + 1) * (2 * jmax + 1));
    oracle_start_norm_table(jmax, start_norm);
    int width = 2 * max_out + 1, wm = 2 * max_m + 1;
    size_t cstride_l = (size_t)npair * 4 * nfreq;                 /* per degree */
    size_t cstride_j = (size_t)(degree + 1) * cstride_l;          /* per interval */
    size_t cstride_lam = (size_t)nint * cstride_j;                /* per band */
    memset(result, 0, sizeof(cf64) * (size_t)nrec * nfreq);

#pragma omp parallel for schedule(dynamic, 1)
    for (int index = 0; index < nrec; ++index) {
        double* wigner = (double*)malloc(sizeof(double) * (degree + 1) * width * (max_m + 1));
        double* detector = (double*)malloc(sizeof(double) * (max_out + 1) * (max_out + 1));
        double* erow = (double*)malloc(sizeof(double) * width * wm);
        cf64* rotated = (cf64*)malloc(sizeof(cf64) * nlam * width * nfreq);
        cf64* channel_sum = (cf64*)malloc(sizeof(cf64) * width * nfreq);
        cf64* scale_phase = (cf64*)malloc(sizeof(cf64) * nlam * nfreq);
        cf64* response = (cf64*)malloc(sizeof(cf64) * nlam * nfreq);
        double cosines[64], sines[64];
        cf64* row = result + (size_t)index * nfreq;
        for (int cell = 0; cell < npoints; ++cell) {
            double dx = receivers[index * 3 + 0] - points[cell * 3 + 0];
            double dy = receivers[index * 3 + 1] - points[cell * 3 + 1];
            double dz = receivers[index * 3 + 2] - points[cell * 3 + 2];
            double radius = sqrt(dx * dx + dy * dy + dz * dz);
            if (radius < radial_low || radius > radial_high) continue;
            double cosine = dz / radius;
            if (cosine > 1.0) cosine = 1.0;
            else if (cosine < -1.0) cosine = -1.0;
            double phi = atan2(dy, dx);
            double transverse = sqrt(dx * dx + dy * dy);
            double sine = transverse / radius;
            double cp = transverse == 0.0 ? 1.0 : dx / transverse;
            double sp = transverse == 0.0 ? 0.0 : dy / transverse;
            for (int n = 0; n <= max_m; ++n) {
                cosines[n] = cos(n * phi);
                sines[n] = sin(n * phi);
            }
            double lx = looks[index * 3 + 0], ly = looks[index * 3 + 1], lz = looks[index * 3 + 2];
            double a1 = cosine * cp * lx + cosine * sp * ly - sine * lz;
            double a2 = -sp * lx + cp * ly;
            double a3 = sine * cp * lx + sine * sp * ly + cosine * lz;
            harmonics_into(a1, a2, a3, max_out, detector);
            wigner_into(degree, max_out, max_m, cosine, start_norm, jmax, wigner);
            double logr = log(radius);
            /* searchsorted(log_r, logr, side='right') - 1, clamped */
            int j = 0;
            while (j < nint + 1 && log_r[j] <= logr) ++j;
            j -= 1;
            if (j < 0) j = 0;
            if (j > nint - 1) j = nint - 1;
            double t = logr - log_r[j];
            for (int lam = 0; lam < nlam; ++lam) {
                double scale = out_scale[(size_t)lam * nrec + index] * exp(-absorption[lam] * radius)
                               / (4 * M_PI * radius * radius);
                for (int w = 0; w < nfreq; ++w) {
                    response[lam * nfreq + w] = 0.0;
                    if (flight) {
                        double angle = omega[w] * radius / velocity[lam];
                        scale_phase[lam * nfreq + w] = scale * (cos(angle) + I * sin(angle));
                    } else {
                        scale_phase[lam * nfreq + w] = scale;
                    }
                }
            }
            for (int ell = 0; ell <= degree; ++ell) {
                rotation_rows(ell, max_out, max_m, wigner, cosines, sines, erow);
                int top_p = ell < max_out ? ell : max_out;
                int keep = ell < max_m ? ell : max_m;
                int64_t centre = starts[ell] + keep;
                memset(rotated, 0, sizeof(cf64) * nlam * width * nfreq);
                for (int a = max_out - top_p; a <= max_out + top_p; ++a) {
                    for (int b = max_m - keep; b <= max_m + keep; ++b) {
                        double factor = erow[a * wm + b];
                        if (factor == 0.0) continue;
                        int64_t column = centre + (b - max_m);
                        for (int lam = 0; lam < nlam; ++lam) {
                            const cf32* ch = channels + (((size_t)lam * npoints + cell) * nchan + column) * nfreq;
                            cf64* rot = rotated + ((size_t)lam * width + a) * nfreq;
                            for (int w = 0; w < nfreq; ++w) rot[w] += factor * (cf64)ch[w];
                        }
                    }
                }
                for (int lam = 0; lam < nlam; ++lam) {
                    memset(channel_sum, 0, sizeof(cf64) * width * nfreq);
                    const cf32* cj = coefficients + lam * cstride_lam + j * cstride_j + ell * cstride_l;
                    for (int p = 0; p < npair; ++p) {
                        int lam_detector = (int)pair_lambda[p];
                        int mu_pair = (int)pair_mu[p];
                        if (mu_pair > top_p) continue;
                        int detector_centre = lam_detector * lam_detector + lam_detector;
                        double positive = alpha[lam_detector] * detector[detector_centre + mu_pair];
                        double negative = mu_pair ? alpha[lam_detector] * detector[detector_centre - mu_pair] : 0.0;
                        if (positive == 0.0 && negative == 0.0) continue;
                        const cf32* c = cj + (size_t)p * 4 * nfreq;
                        for (int w = 0; w < nfreq; ++w) {
                            cf64 c0 = c[w], c1 = c[nfreq + w], c2 = c[2 * nfreq + w], c3 = c[3 * nfreq + w];
                            cf64 radial = ((c0 * t + c1) * t + c2) * t + c3;
                            if (negative != 0.0) channel_sum[(max_out - mu_pair) * nfreq + w] += negative * radial;
                            if (positive != 0.0) channel_sum[(max_out + mu_pair) * nfreq + w] += positive * radial;
                        }
                    }
                    for (int a = max_out - top_p; a <= max_out + top_p; ++a) {
                        for (int w = 0; w < nfreq; ++w) {
                            cf64 value = rotated[((size_t)lam * width + a) * nfreq + w];
                            if (value == 0.0) continue;
                            response[lam * nfreq + w] += value * channel_sum[a * nfreq + w];
                        }
                    }
                }
            }
            for (int lam = 0; lam < nlam; ++lam)
                for (int w = 0; w < nfreq; ++w)
                    row[w] += response[lam * nfreq + w] * scale_phase[lam * nfreq + w];
        }
        free(wigner); free(detector); free(erow); free(rotated);
        free(channel_sum); free(scale_phase); free(response);
    }
    free(start_norm);
}
