/*
 * Oracle of stage (a): order-0 (direct) light to the first glass entry, FP64.
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 *
 * Quantity (the first-entry integral of PreparedBallistic,
 * _surface_ballistic.py:20-353; its integrand is written out by the reference's
 * direct-ray evaluator _surface_ballistic_rays.py:40-156 and by
 * tests/test_ballistic_first_entry.py:77-123):
 *
 *   q0[om] = area/(pi R^2) sum_i int ds int dlambda S_i(lambda)
 *            (1/2pi) int dphi A(d.look) exp(-mu_ext(lambda) d_entry)
 *   S_i = (C0_i + C2_i/n^2) / (L_i lambda^2) * eta(lambda)
 *
 * over photons emitted on the Cherenkov cone (cos theta = 1/(beta n)) that
 * hit the OM sphere; bins split it by arrival time
 *   t = t0 + tau s + d_entry n_g / c - origin      (half-open bins).
 *
 * The reference reaches it as a control-variate sum (leading finite-cone
 * term + residual + omission bounds) converged to 2e-6 relative + 1e-10 pe
 * per OM. This oracle evaluates the same integral directly to ~1e-10:
 *   - candidate pairs: interval_candidates (_ballistic_interval_numba.py:
 *     450-519) with _cone_can_reach_sphere and _support (:152-291), the glass
 *     check of _surface_ballistic.py:48-53;
 *   - regular pairs (rho > R, no constant-phase panel; _surface_ballistic.py:
 *     54-56): emission point s outer, cone angle theta middle, azimuth inner.
 *     At fixed s the OM is a cap of angular radius eps = asin(R/D) around the
 *     direction psi of its centre; d lambda = |d lambda/d theta| d theta per
 *     optical panel. theta = psi + eps sin(t) and phi = half sin(eta) remove
 *     the square-root edges of the cap; s is split where a panel edge is
 *     tangent to the cap (s = z - rho cot(theta_k) -+ R / sin(theta_k));
 *     bins split phi at the exact edge crossings and theta where an edge
 *     crosses the cap's centre meridian;
 *   - other pairs: the reference's direct-ray evaluator restated.
 * Each pair is integrated with doubling orders (4 ... 64 nodes per
 * dimension) until its charge changes by less than rel_tol and its bin L1
 * by less than bin_tol, both relative to its charge. Charge converges
 * geometrically; bins converge algebraically once the emission time across
 * an element is comparable with a bin, hence the separate tolerance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define C_LIGHT 0.299792458

typedef struct {
    int npanel;
    const double* edges;   /* (npanel+1) nm */
    const double* phase;   /* (npanel+1) */
    const double* group;   /* (npanel+1) */
    const double* ext;     /* (npanel+1) 1/m */
    const double* eff;     /* (npanel, effcols) */
    int effcols;
} Optics;

typedef struct {
    double z, rho, axial, across, side, L, c0, c2, beta, t0, tau, R, norm;
} PairGeom;

/* ---------------------------------------------------------------- helpers */

static double acceptance(const double* alpha, int n, double x) {
    /* _acceptance (_ballistic_interval_numba.py:21-37) */
    double result = alpha[0] / (4 * M_PI);
    double previous = 1.0, current = x;
    if (n > 1) result += 3 * alpha[1] * current / (4 * M_PI);
    for (int ell = 2; ell < n; ++ell) {
        double value = ((2 * ell - 1) * x * current - (ell - 1) * previous) / ell;
        result += (2 * ell + 1) * alpha[ell] * value / (4 * M_PI);
        previous = current;
        current = value;
    }
    return result > 0 ? result : 0.0;
}

static double efficiency(const Optics* o, int k, double q) {
    const double* c = o->eff + (size_t)k * o->effcols;
    double eta = ((c[3] * q + c[2]) * q + c[1]) * q + c[0];
    if (o->effcols > 4) eta *= exp(c[4] + q * c[5]);
    return eta;
}

static void gauss_legendre(int n, double* x, double* w) {
    /* Newton on P_n; nodes ascending on [-1, 1]. */
    for (int i = 0; i < n; ++i) {
        double r = cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = r;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2 * k - 1) * r * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = r; p0 = 1.0; }
            double dp = n * (r * p1 - p0) / (r * r - 1);
            double dr = p1 / dp;
            r -= dr;
            if (fabs(dr) < 1e-16) break;
        }
        double p0 = 1.0, p1 = r;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2 * k - 1) * r * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) { p1 = r; p0 = 1.0; }
        double dp = n * (r * p1 - p0) / (r * r - 1);
        x[n - 1 - i] = r;
        w[n - 1 - i] = 2.0 / ((1 - r * r) * dp * dp);
    }
}

static int cmp_double(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* bins own their left edge: index of the bin holding t, or -1 */
static int bin_of(const double* edges, int nbins, double t) {
    if (!(t >= edges[0]) || !(t < edges[nbins])) return -1;
    int lo = 0, hi = nbins;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (edges[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------ candidate pairs */

/* _cone_can_reach_sphere (_ballistic_interval_numba.py:226-252) */
static int cone_can_reach_sphere(double z, double rho, double length, double beta, double radius,
                                 double phase_min, double phase_max) {
    if (z <= -radius || beta * phase_max <= 1.) return 0;
    double c_hi = 1. / (beta * phase_max);
    double s_hi = sqrt(fmax(0., (1. - c_hi) * (1. + c_hi)));
    double c_lo = fmin(1., 1. / (beta * phase_min));
    double s_lo = sqrt(fmax(0., (1. - c_lo) * (1. + c_lo)));
    double q_hi = rho * c_hi - z * s_hi;
    double minimum = z >= 0. ? q_hi : fmin(q_hi, rho * c_lo - z * s_lo);
    double rounding = 16 * 2.220446049250313e-16 * (fabs(z) + rho + length + radius);
    if (minimum > radius + rounding) return 0;
    double end_z = z - length;
    if (end_z >= 0. && rho * c_lo - end_z * s_lo < -radius - rounding) return 0;
    return 1;
}

/* _support (_ballistic_interval_numba.py:152-186), radius > 0 */
static int support(double z, double rho, double length, double beta, double radius,
                   double* nmin, double* nmax) {
    const double quarter = 0.5 * M_PI;
    double psi0 = atan2(rho, z), psi1 = atan2(rho, z - length);
    double a0 = asin(fmin(1., radius / hypot(z, rho)));
    double a1 = asin(fmin(1., radius / hypot(z - length, rho)));
    double low = psi0 - a0, high = psi1 + a1;
    if (low >= quarter) return 0;
    *nmin = low <= 0. ? (1. + 1e-9) / beta : fmax((1. + 1e-9) / beta, 1. / (beta * cos(low)));
    *nmax = high >= quarter ? INFINITY : 1. / (beta * cos(high));
    return 1;
}

/* interval_candidates(include_surface_axis=True) for spheres with infinite
 * reach. Receiver-major, element order. Writes up to capacity pairs; returns
 * the count (may exceed capacity: call again with a larger buffer), or -1 if
 * an emitting segment intersects the glass. */
int64_t oracle_order0_candidates(const double* positions, const double* radii, int nom,
                                 const double* start, const double* direction,
                                 const double* length, const double* c0, const double* beta,
                                 int nel, double phase_min, double phase_max,
                                 int64_t* pair_om, int64_t* pair_el, int64_t capacity) {
    int64_t count = 0;
    for (int om = 0; om < nom; ++om) {
        double R = radii[om];
        for (int i = 0; i < nel; ++i) {
            if (c0[i] == 0. || beta[i] * phase_max <= 1.) continue;
            double dx = positions[3 * om] - start[3 * i];
            double dy = positions[3 * om + 1] - start[3 * i + 1];
            double dz = positions[3 * om + 2] - start[3 * i + 2];
            double ux = direction[3 * i], uy = direction[3 * i + 1], uz = direction[3 * i + 2];
            double z = dx * ux + dy * uy + dz * uz;
            if (z <= 0. && z <= -R) continue;
            double cx = dy * uz - dz * uy, cy = dz * ux - dx * uz, cz = dx * uy - dy * ux;
            double rho = sqrt(cx * cx + cy * cy + cz * cz);
            if (rho != 0.) {
                if (!cone_can_reach_sphere(z, rho, length[i], beta[i], R, phase_min, phase_max))
                    continue;
                double nmin, nmax;
                if (!support(z, rho, length[i], beta[i], R, &nmin, &nmax)) continue;
                if (nmin > phase_max || nmax <= phase_min) continue;
            }
            /* emitting segment intersects the glass (_surface_ballistic.py:48-53) */
            double along = fmax(fmax(-z, z - length[i]), 0.);
            if (hypot(rho, along) <= R) return -1;
            if (count < capacity) {
                pair_om[count] = om;
                pair_el[count] = i;
            }
            ++count;
        }
    }
    return count;
}

/* ---------------------------------------------------- regular pairs: cap */

typedef struct {
    const Optics* o;
    const PairGeom* g;
    const double* alpha;
    int nalpha;
    const double* edges;
    int nbins;
    double origin;
    int ns, nt, ne;
    const double *gxs, *gws, *gxt, *gwt, *gxe, *gwe;
    double theta_edge[4096];
    int valid_edge[4096];
} CapCtx;

/* Entry distance and its haversine form at fixed (s, theta): everything in
 * terms of the cap offset delta = theta - psi so nothing cancels. */
typedef struct {
    double D, D2R2, psi, sinpsi, hav_eps, eps;
} SGeom;

static double hav(double x) { double s = sin(0.5 * x); return s * s; }

/* d_entry for a direction at cap offset delta (hav_delta) and azimuth phi */
static double entry_distance(const SGeom* sg, double sin_theta, double hav_delta, double hav_phi) {
    double hav_gamma = hav_delta + sin_theta * sg->sinpsi * hav_phi;
    double inside = (sg->hav_eps - hav_gamma) * (1.0 - sg->hav_eps - hav_gamma);
    if (inside < 0.) inside = 0.;
    return sg->D2R2 / (sg->D * (1.0 - 2.0 * hav_gamma) + 2.0 * sg->D * sqrt(inside));
}

/* hav(gamma) of the direction that enters at distance d: triangle
 * (emission, centre, entry point), R^2 = D^2 + d^2 - 2 D d cos(gamma). */
static double hav_gamma_of_distance(const SGeom* sg, double R, double d) {
    return (R - sg->D + d) * (R + sg->D - d) / (4.0 * sg->D * d);
}

/* Integral over one theta-piece [ta, tb] (t-variable) of panel k at emission
 * point s; adds charge (returned) and bins. */
static double cap_piece(CapCtx* c, const SGeom* sg, int k, double ta, double tb, double base_time,
                        double ds_weight, double* bins, int split_time) {
    const Optics* o = c->o;
    const PairGeom* g = c->g;
    double left = o->edges[k], width = o->edges[k + 1] - left;
    double n0 = o->phase[k], dn = o->phase[k + 1] - n0;
    double g0 = o->group[k], dg = o->group[k + 1] - g0;
    double e0 = o->ext[k], de = o->ext[k + 1] - e0;
    double total = 0.;
    /* time split of [ta, tb]: where a bin edge crosses the centre meridian
     * (phi = 0, found by sampling and bisection) or the rim of the cap (there
     * d_entry = D cos(eps) for every theta, so the crossing is where the
     * group index takes one value: closed form in the panel). */
    double cuts[512];
    int ncut = 0;
    cuts[ncut++] = ta;
    if (split_time && bins) {
        const int M = 48;
        double tv[49], tt[49];
        for (int m = 0; m <= M; ++m) {
            double t = ta + (tb - ta) * m / M;
            double delta = sg->eps * sin(t);
            double theta = sg->psi + delta;
            double nn = 1.0 / (g->beta * cos(theta));
            double q = (nn - n0) / dn;
            double d = entry_distance(sg, sin(theta), hav(delta), 0.0);
            tt[m] = t;
            tv[m] = base_time + d * (g0 + q * dg) / C_LIGHT;
        }
        double rim = sg->D * cos(sg->eps);
        /* edges that can matter: inside the sampled meridian times or the rim times */
        double lo = tv[0], hi = tv[0];
        for (int m = 1; m <= M; ++m) { lo = fmin(lo, tv[m]); hi = fmax(hi, tv[m]); }
        double rim_a = base_time + rim * g0 / C_LIGHT, rim_b = base_time + rim * (g0 + dg) / C_LIGHT;
        lo = fmin(lo, fmin(rim_a, rim_b));
        hi = fmax(hi, fmax(rim_a, rim_b));
        int e0i = 0, e1i = c->nbins + 1;
        while (e0i < c->nbins + 1 && c->edges[e0i] <= lo) ++e0i;
        while (e1i > e0i && c->edges[e1i - 1] >= hi) --e1i;
        for (int e = e0i; e < e1i && ncut < 500; ++e) {
            double T = c->edges[e];
            for (int m = 0; m < M && ncut < 500; ++m) {
                double a_lo = fmin(tv[m], tv[m + 1]), a_hi = fmax(tv[m], tv[m + 1]);
                if (!(T > a_lo && T < a_hi)) continue;
                double a = tt[m], b = tt[m + 1], fa = tv[m] - T;
                for (int it = 0; it < 80; ++it) {
                    double mid = 0.5 * (a + b);
                    if (mid == a || mid == b) break;
                    double delta = sg->eps * sin(mid);
                    double theta = sg->psi + delta;
                    double nn = 1.0 / (g->beta * cos(theta));
                    double q = (nn - n0) / dn;
                    double d = entry_distance(sg, sin(theta), hav(delta), 0.0);
                    double fm = base_time + d * (g0 + q * dg) / C_LIGHT - T;
                    if ((fa < 0.) == (fm < 0.)) { a = mid; fa = fm; } else b = mid;
                }
                cuts[ncut++] = 0.5 * (a + b);
            }
            if (dg != 0. && T > base_time) {
                double target = (T - base_time) * C_LIGHT / rim;
                double q = (target - g0) / dg;
                if (q > 0. && q < 1.) {
                    double nn = n0 + q * dn;
                    if (g->beta * nn > 1.) {
                        double x = (acos(1.0 / (g->beta * nn)) - sg->psi) / sg->eps;
                        if (x > -1. && x < 1.) {
                            double t = asin(x);
                            if (t > ta && t < tb) cuts[ncut++] = t;
                        }
                    }
                }
            }
        }
    }
    cuts[ncut++] = tb;
    qsort(cuts, ncut, sizeof(double), cmp_double);
    for (int piece = 0; piece + 1 < ncut; ++piece) {
        double pa = cuts[piece], pb = cuts[piece + 1];
        if (!(pb > pa)) continue;
        for (int it = 0; it < c->nt; ++it) {
            double xm = 0.5 * M_PI * c->gxt[it];
            double t = 0.5 * (pa + pb) + 0.5 * (pb - pa) * sin(xm);
            double wt = 0.5 * (pb - pa) * cos(xm) * 0.5 * M_PI * c->gwt[it];
            double st = sin(t), ct = cos(t);
            double delta = sg->eps * st;
            double theta = sg->psi + delta;
            double sin_theta = sin(theta), cos_theta = cos(theta);
            double nn = 1.0 / (g->beta * cos_theta);
            double q = (nn - n0) / dn;
            if (q < -1e-9 || q > 1 + 1e-9) continue;
            double lam = left + width * q;
            double eta = efficiency(o, k, q);
            double spectral = (g->c0 + g->c2 / (nn * nn)) / g->L / (lam * lam) * eta;
            /* |d lambda / d theta| = beta n^2 sin(theta) width / |dn| */
            double jac = g->beta * nn * nn * sin_theta * width / fabs(dn);
            double ext = e0 + q * de;
            double group = g0 + q * dg;
            /* half-width of the azimuth range inside the cap */
            double H = sin(0.5 * sg->eps * (1 - st)) * sin(0.5 * sg->eps * (1 + st));
            double hh = H / (sin_theta * sg->sinpsi);
            if (hh <= 0.) continue;
            if (hh > 1.) hh = 1.;
            double half = 2.0 * asin(sqrt(hh));
            double hav_delta = hav(delta);
            double weight = ds_weight * wt * sg->eps * ct * spectral * jac * g->norm;
            /* azimuth pieces: [0, pi/2] in eta, split at time-edge crossings */
            double ecut[512];
            int necut = 0;
            ecut[necut++] = 0.;
            if (bins) {
                double d0 = entry_distance(sg, sin_theta, hav_delta, 0.);
                double d1 = entry_distance(sg, sin_theta, hav_delta, hav(half));
                double t0 = base_time + d0 * group / C_LIGHT;
                double t1 = base_time + d1 * group / C_LIGHT;
                int e = 0;
                {
                    int lo_i = 0, hi_i = c->nbins + 1;
                    while (lo_i < hi_i) { int mid = (lo_i + hi_i) / 2; if (c->edges[mid] <= t0) lo_i = mid + 1; else hi_i = mid; }
                    e = lo_i;
                }
                for (; e <= c->nbins && necut < 510; ++e) {
                    double T = c->edges[e];
                    if (!(T < t1)) break;
                    double d = (T - base_time) * C_LIGHT / group;
                    double hg = hav_gamma_of_distance(sg, c->g->R, d);
                    double hp = (hg - hav_delta) / (sin_theta * sg->sinpsi);
                    if (hp <= 0. || hp >= 1.) continue;
                    double phi = 2.0 * asin(sqrt(hp));
                    double x = phi / half;
                    if (x <= 0. || x >= 1.) continue;
                    ecut[necut++] = asin(x);
                }
            }
            ecut[necut++] = 0.5 * M_PI;
            if (necut > 2) qsort(ecut, necut, sizeof(double), cmp_double);
            for (int ep = 0; ep + 1 < necut; ++ep) {
                double ea = ecut[ep], eb = ecut[ep + 1];
                if (!(eb > ea)) continue;
                double integral = 0.;
                for (int ie = 0; ie < c->ne; ++ie) {
                    double e = 0.5 * (ea + eb) + 0.5 * (eb - ea) * c->gxe[ie];
                    double phi = half * sin(e);
                    double dphi = half * cos(e) * 0.5 * (eb - ea) * c->gwe[ie];
                    double cp = cos(phi), sp = sin(phi);
                    double d = entry_distance(sg, sin_theta, hav_delta, hav(phi));
                    double acc = acceptance(c->alpha, c->nalpha,
                                            cos_theta * g->axial + sin_theta * (cp * g->across + sp * g->side))
                               + acceptance(c->alpha, c->nalpha,
                                            cos_theta * g->axial + sin_theta * (cp * g->across - sp * g->side));
                    integral += dphi * acc * exp(-ext * d);
                }
                double value = weight * integral / (2 * M_PI);
                total += value;
                if (bins) {
                    double phim = half * sin(0.5 * (ea + eb));
                    double d = entry_distance(sg, sin_theta, hav_delta, hav(phim));
                    int b = bin_of(c->edges, c->nbins, base_time + d * group / C_LIGHT);
                    if (b >= 0) bins[b] += value;
                }
            }
        }
    }
    return total;
}

static double cap_pair(CapCtx* c, double* bins, int split_time) {
    const PairGeom* g = c->g;
    const Optics* o = c->o;
    /* panel edges as cone angles */
    for (int k = 0; k <= o->npanel; ++k) {
        double bn = g->beta * o->phase[k];
        c->valid_edge[k] = bn > 1.;
        c->theta_edge[k] = bn > 1. ? acos(1.0 / bn) : 0.0;
    }
    /* s pieces: [0, L] split where a panel edge is tangent to the cap */
    double scuts[8200];
    int nsc = 0;
    scuts[nsc++] = 0.;
    for (int k = 0; k <= o->npanel; ++k) {
        if (!c->valid_edge[k]) continue;
        double th = c->theta_edge[k];
        double centre = g->z - g->rho * cos(th) / sin(th);
        double spread = g->R / sin(th);
        double r1 = centre - spread, r2 = centre + spread;
        if (r1 > 0. && r1 < g->L) scuts[nsc++] = r1;
        if (r2 > 0. && r2 < g->L) scuts[nsc++] = r2;
    }
    scuts[nsc++] = g->L;
    qsort(scuts, nsc, sizeof(double), cmp_double);
    double total = 0.;
    for (int sp = 0; sp + 1 < nsc; ++sp) {
        double sa = scuts[sp], sb = scuts[sp + 1];
        if (!(sb > sa)) continue;
        for (int is = 0; is < c->ns; ++is) {
            double x = M_PI / 2 * c->gxs[is];
            double s = 0.5 * (sa + sb) + 0.5 * (sb - sa) * sin(x);
            double ds = 0.5 * (sb - sa) * cos(x) * M_PI / 2 * c->gws[is];
            SGeom sg;
            double along = g->z - s;
            sg.D = sqrt(g->rho * g->rho + along * along);
            sg.D2R2 = (sg.D - g->R) * (sg.D + g->R);
            sg.psi = atan2(g->rho, along);
            sg.sinpsi = g->rho / sg.D;
            sg.eps = asin(g->R / sg.D);
            sg.hav_eps = hav(sg.eps);
            double base_time = g->t0 + g->tau * s - c->origin;
            for (int k = 0; k < o->npanel; ++k) {
                if (!c->valid_edge[k] && !c->valid_edge[k + 1]) continue;
                double th0 = c->theta_edge[k], th1 = c->theta_edge[k + 1];
                double lo = fmin(th0, th1), hi = fmax(th0, th1);
                double xa = (lo - sg.psi) / sg.eps, xb = (hi - sg.psi) / sg.eps;
                if (xb <= -1. || xa >= 1.) continue;
                double ta = xa <= -1. ? -M_PI / 2 : asin(xa);
                double tb = xb >= 1. ? M_PI / 2 : asin(xb);
                if (!(tb > ta)) continue;
                total += cap_piece(c, &sg, k, ta, tb, base_time, ds, bins, split_time);
            }
        }
    }
    return total;
}

/* ------------------------------------------- other pairs: direct rays */

/* _quadratic_cuts (_surface_ballistic_rays.py:15-31) */
static int quadratic_cuts(double* cuts, int count, double a, double b, double c, double low,
                          double high) {
    if (a == 0.) {
        if (b != 0.) {
            double root = -c / b;
            if (low < root && root < high) cuts[count++] = root;
        }
        return count;
    }
    double disc = b * b - 4 * a * c;
    if (disc < 0.) return count;
    double q = -0.5 * (b + copysign(sqrt(disc), b));
    if (q == 0.) {
        double root = -b / (2 * a);
        if (low < root && root < high) cuts[count++] = root;
    } else {
        double r1 = q / a, r2 = c / q;
        if (low < r1 && r1 < high) cuts[count++] = r1;
        if (low < r2 && r2 < high) cuts[count++] = r2;
    }
    return count;
}

static double entry_b(double B, double dot) { return B / (dot + sqrt(fmax(0., dot * dot - B))); }

static int searchsorted_right(const double* a, int n, double v) {
    int lo = 0, hi = n;
    while (lo < hi) { int mid = (lo + hi) / 2; if (a[mid] <= v) lo = mid + 1; else hi = mid; }
    return lo;
}
static int searchsorted_left(const double* a, int n, double v) {
    int lo = 0, hi = n;
    while (lo < hi) { int mid = (lo + hi) / 2; if (a[mid] < v) lo = mid + 1; else hi = mid; }
    return lo;
}

/* evaluate (_surface_ballistic_rays.py:40-156) for one pair */
static double rays_pair(const Optics* o, const PairGeom* g, const double* alpha, int nalpha,
                        const double* gx, const double* gw, int n, double origin,
                        const double* edges, int nedges, double* bins) {
    double phase_min = o->phase[0], phase_max = o->phase[0];
    for (int k = 0; k <= o->npanel; ++k) {
        phase_min = fmin(phase_min, o->phase[k]);
        phase_max = fmax(phase_max, o->phase[k]);
    }
    /* _support for this pair */
    double nmin, nmax;
    double cut0 = INFINITY, cut1 = INFINITY;
    {
        const double quarter = 0.5 * M_PI;
        double psi0 = atan2(g->rho, g->z), psi1 = atan2(g->rho, g->z - g->L);
        double a0 = asin(fmin(1., g->R / hypot(g->z, g->rho)));
        double a1 = asin(fmin(1., g->R / hypot(g->z - g->L, g->rho)));
        double low = psi0 - a0, high = psi1 + a1;
        if (low >= quarter) return 0.;
        nmin = low <= 0. ? (1. + 1e-9) / g->beta : fmax((1. + 1e-9) / g->beta, 1. / (g->beta * cos(low)));
        nmax = high >= quarter ? INFINITY : 1. / (g->beta * cos(high));
        double inner0 = psi0 + a0, inner1 = psi1 - a1;
        cut0 = inner0 >= quarter ? INFINITY : 1. / (g->beta * cos(inner0));
        cut1 = (inner1 <= 0. || inner1 >= quarter) ? INFINITY : 1. / (g->beta * cos(inner1));
    }
    if (nmin > phase_max || nmax <= phase_min) return 0.;
    double total = 0.;
    double velocity_c = C_LIGHT;
    double phicuts[1100];
    double scuts[6 * 1100 + 2];
    for (int panel = 0; panel < o->npanel; ++panel) {
        double left = o->edges[panel], width = o->edges[panel + 1] - left;
        double n0 = o->phase[panel], dn = o->phase[panel + 1] - n0;
        double a = left, b = left + width;
        if (dn == 0.) {
            if (n0 < nmin || n0 >= nmax) continue;
        } else {
            double x0 = (nmin - n0) / dn, x1 = (nmax - n0) / dn;
            a = left + width * fmax(0., fmin(x0, x1));
            b = left + width * fmin(1., fmax(x0, x1));
        }
        if (a >= b) continue;
        double cuts[7];
        int count = 0;
        cuts[count++] = a;
        cuts[count++] = b;
        if (dn != 0.) {
            double cv[2] = {cut0, cut1};
            for (int j = 0; j < 2; ++j) {
                if (isfinite(cv[j])) {
                    double lam = left + width * (cv[j] - n0) / dn;
                    if (a < lam && lam < b) cuts[count++] = lam;
                }
            }
            qsort(cuts, count, sizeof(double), cmp_double);
        }
        for (int piece = 0; piece + 1 < count; ++piece) {
            double low = cuts[piece], high = cuts[piece + 1];
            for (int il = 0; il < n; ++il) {
                double lam = (low + high) / 2 + (high - low) / 2 * gx[il];
                double q = (lam - left) / width, nn = n0 + dn * q, bn = g->beta * nn;
                if (bn <= 1.) continue;
                double cosine = 1 / bn, sine = sqrt(1 - cosine * cosine);
                double centre = g->z - g->rho * cosine / sine, spread = g->R / sine;
                double slo = fmax(0., centre - spread), shi = fmin(g->L, centre + spread);
                if (slo >= shi) continue;
                double absorption = o->ext[panel] + q * (o->ext[panel + 1] - o->ext[panel]);
                double group = o->group[panel] + q * (o->group[panel + 1] - o->group[panel]);
                double velocity = velocity_c / group;
                double eta = efficiency(o, panel, q);
                double spectral = (g->c0 + g->c2 / (nn * nn)) / g->L / (lam * lam) * eta;
                double scale = (high - low) / 2 * gw[il] * spectral * g->norm;
                int ns = 0;
                scuts[ns++] = slo;
                scuts[ns++] = shi;
                if (bins) {
                    double rmax = sqrt(g->rho * g->rho + fmax((g->z - slo) * (g->z - slo), (g->z - shi) * (g->z - shi)));
                    double mn = (slo < g->z && g->z < shi) ? 0. : fmin((g->z - slo) * (g->z - slo), (g->z - shi) * (g->z - shi));
                    double rmin = sqrt(g->rho * g->rho + mn);
                    double earliest = fmin(g->t0 + g->tau * slo, g->t0 + g->tau * shi) + (rmin - g->R) / velocity - origin;
                    double latest = fmax(g->t0 + g->tau * slo, g->t0 + g->tau * shi) + rmax / velocity - origin;
                    int first_edge = searchsorted_right(edges, nedges, earliest);
                    int last_edge = searchsorted_left(edges, nedges, latest);
                    for (int edge = first_edge; edge < last_edge && ns < 6 * 1100; ++edge) {
                        double A = (origin + edges[edge] - g->t0) * velocity, Bq = g->tau * velocity;
                        for (int sgn = -1; sgn <= 1; sgn += 2) {
                            double Dd = cosine * g->z + sgn * sine * g->rho;
                            ns = quadratic_cuts(scuts, ns, 1 + Bq * Bq - 2 * Bq * cosine,
                                                -2 * g->z - 2 * A * Bq + 2 * (A * cosine + Bq * Dd),
                                                g->z * g->z + g->rho * g->rho - g->R * g->R + A * A - 2 * A * Dd, slo, shi);
                        }
                        ns = quadratic_cuts(scuts, ns, Bq * Bq - 1., -2 * A * Bq + 2 * g->z,
                                            A * A - g->z * g->z - g->rho * g->rho + g->R * g->R, slo, shi);
                    }
                }
                qsort(scuts, ns, sizeof(double), cmp_double);
                for (int spc = 0; spc + 1 < ns; ++spc) {
                    double sa = scuts[spc], sb = scuts[spc + 1];
                    if (sa >= sb) continue;
                    for (int j = 0; j < n; ++j) {
                        double psi = M_PI / 2 * gx[j];
                        double s = (sa + sb) / 2 + (sb - sa) / 2 * sin(psi);
                        double ds = (sb - sa) / 2 * cos(psi) * M_PI / 2 * gw[j];
                        double along = g->z - s, B = g->rho * g->rho + along * along - g->R * g->R;
                        double half;
                        if (g->rho > 0.)
                            half = acos(fmin(1., fmax(-1., (sqrt(B) - cosine * along) / (sine * g->rho))));
                        else
                            half = cosine * along >= sqrt(B) ? M_PI : 0.;
                        if (half == 0.) continue;
                        double emission = g->t0 + g->tau * s - origin;
                        int nphi = 0;
                        phicuts[nphi++] = 0.;
                        if (bins && g->rho > 0.) {
                            double tmin = emission + entry_b(B, cosine * along + sine * g->rho) / velocity;
                            double tmax = emission + entry_b(B, cosine * along + sine * g->rho * cos(half)) / velocity;
                            int first = searchsorted_right(edges, nedges, tmin);
                            int last = searchsorted_left(edges, nedges, tmax);
                            for (int edge = first; edge < last && nphi < 1090; ++edge) {
                                double distance = (edges[edge] - emission) * velocity;
                                double dot = (B + distance * distance) / (2 * distance);
                                double phi = acos(fmin(1., fmax(-1., (dot - cosine * along) / (sine * g->rho))));
                                phicuts[nphi++] = asin(fmin(1., phi / half));
                            }
                        }
                        phicuts[nphi++] = M_PI / 2;
                        for (int pp = 0; pp + 1 < nphi; ++pp) {
                            double pa = phicuts[pp], pb = phicuts[pp + 1];
                            if (pa >= pb) continue;
                            double integral = 0.;
                            for (int jp = 0; jp < n; ++jp) {
                                double eta_phi = (pa + pb) / 2 + (pb - pa) / 2 * gx[jp];
                                double phi = half * sin(eta_phi);
                                double dphi = half * cos(eta_phi) * (pb - pa) / 2 * gw[jp];
                                double cp = cos(phi), ss = sin(phi);
                                double distance = entry_b(B, cosine * along + sine * g->rho * cp);
                                double acc = acceptance(alpha, nalpha, cosine * g->axial + sine * (cp * g->across + ss * g->side))
                                           + acceptance(alpha, nalpha, cosine * g->axial + sine * (cp * g->across - ss * g->side));
                                integral += dphi * acc * exp(-absorption * distance) / (2 * M_PI);
                            }
                            double value = scale * ds * integral;
                            total += value;
                            if (bins) {
                                double phi_mid = half * sin((pa + pb) / 2);
                                double arrival = emission + entry_b(B, cosine * along + sine * g->rho * cos(phi_mid)) / velocity;
                                int slot = searchsorted_right(edges, nedges, arrival) - 1;
                                if (slot >= 0 && slot < nedges - 1) bins[slot] += value;
                            }
                        }
                    }
                }
            }
        }
    }
    return total;
}

/* ------------------------------------------------------------ driver */

static void pair_geometry(PairGeom* g, const double* p, const double* look, double area, double R,
                          const double* x0, const double* u, double L, double c0, double c2,
                          double beta, double t0, double t1) {
    double w[3] = {p[0] - x0[0], p[1] - x0[1], p[2] - x0[2]};
    double z = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
    double tr[3] = {w[0] - z * u[0], w[1] - z * u[1], w[2] - z * u[2]};
    double rho = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]);
    double e1[3];
    if (rho > 0.) {
        e1[0] = tr[0] / rho; e1[1] = tr[1] / rho; e1[2] = tr[2] / rho;
    } else {
        double aux[3] = {1., 0., 0.};
        if (fabs(u[0]) >= .9) { aux[0] = 0.; aux[1] = 1.; }
        e1[0] = u[1] * aux[2] - u[2] * aux[1];
        e1[1] = u[2] * aux[0] - u[0] * aux[2];
        e1[2] = u[0] * aux[1] - u[1] * aux[0];
        double nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        e1[0] /= nrm; e1[1] /= nrm; e1[2] /= nrm;
    }
    double e2[3] = {u[1] * e1[2] - u[2] * e1[1], u[2] * e1[0] - u[0] * e1[2], u[0] * e1[1] - u[1] * e1[0]};
    g->z = z;
    g->rho = rho;
    g->axial = u[0] * look[0] + u[1] * look[1] + u[2] * look[2];
    g->across = e1[0] * look[0] + e1[1] * look[1] + e1[2] * look[2];
    g->side = e2[0] * look[0] + e2[1] * look[1] + e2[2] * look[2];
    g->L = L;
    g->c0 = c0;
    g->c2 = c2;
    g->beta = beta;
    g->t0 = t0;
    g->tau = (t1 - t0) / L;
    g->R = R;
    g->norm = area / (M_PI * R * R);
}

/*
 * Charge (nom) and optional bins (nom, nbins) of the listed pairs (receiver-
 * major). origins (nom) and edges (nbins+1) only with bins. rel_tol: per-pair
 * convergence target. Returns 0, or 1 if a pair failed to converge.
 */
int oracle_order0(const int64_t* pair_om, const int64_t* pair_el, int64_t npairs,
                  const double* positions, const double* looks, const double* areas,
                  const double* radii, int nom, const double* start, const double* direction,
                  const double* length, const double* c0, const double* c2, const double* beta,
                  const double* t0, const double* t1, int npanel, const double* panel_edges,
                  const double* phase, const double* group, const double* ext, const double* eff,
                  int effcols, const double* alpha, int nalpha, const double* origins,
                  const double* edges, int nbins, double rel_tol, double abs_tol,
                  double bin_tol, int force_rays, double* charge, double* bins,
                  double* bins_change) {
    Optics o = {npanel, panel_edges, phase, group, ext, eff, effcols};
    int regular_phase = 1;
    for (int k = 0; k < npanel; ++k)
        if (phase[k + 1] == phase[k]) regular_phase = 0;
    memset(charge, 0, sizeof(double) * nom);
    if (bins) memset(bins, 0, sizeof(double) * (size_t)nom * nbins);
    if (bins_change) memset(bins_change, 0, sizeof(double) * nom);
    /* row boundaries */
    int64_t* first = (int64_t*)calloc((size_t)nom + 1, sizeof(int64_t));
    for (int64_t k = 0; k < npairs; ++k) first[pair_om[k] + 1]++;
    for (int om = 0; om < nom; ++om) first[om + 1] += first[om];
    int failed = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : failed)
    for (int om = 0; om < nom; ++om) {
        double* pb = bins ? (double*)calloc(nbins, sizeof(double)) : NULL;
        double* pb2 = bins ? (double*)calloc(nbins, sizeof(double)) : NULL;
        double* rowb = bins ? bins + (size_t)om * nbins : NULL;
        CapCtx* c = (CapCtx*)malloc(sizeof(CapCtx));
        double gx[6][128], gw[6][128];
        int orders[6] = {4, 8, 16, 32, 64, 128};
        for (int m = 0; m < 6; ++m) gauss_legendre(orders[m], gx[m], gw[m]);
        double origin = origins ? origins[om] : 0.;
        for (int64_t k = first[om]; k < first[om + 1]; ++k) {
            int64_t i = pair_el[k];
            PairGeom g;
            pair_geometry(&g, positions + 3 * om, looks + 3 * om, areas[om], radii[om],
                          start + 3 * i, direction + 3 * i, length[i], c0[i], c2[i], beta[i],
                          t0[i], t1[i]);
            int regular = regular_phase && g.rho > g.R && !force_rays;
            double previous = NAN, value = 0., last_l1 = 0.;
            int converged = 0, charge_converged = 0;
            for (int m = 0; m < (regular ? 5 : 5); ++m) {
                if (pb) memset(pb, 0, sizeof(double) * nbins);
                if (regular) {
                    c->o = &o; c->g = &g; c->alpha = alpha; c->nalpha = nalpha;
                    c->edges = edges; c->nbins = nbins; c->origin = origin;
                    int lvl = m;  /* 4, 8, 16, 32, 64 */
                    c->ns = orders[lvl]; c->nt = orders[lvl]; c->ne = orders[lvl];
                    c->gxs = gx[lvl]; c->gws = gw[lvl]; c->gxt = gx[lvl]; c->gwt = gw[lvl];
                    c->gxe = gx[lvl]; c->gwe = gw[lvl];
                    value = cap_pair(c, pb, 1);
                } else {
                    int lvl = m + 1;  /* 8, 16, 32, 64, 128 */
                    value = rays_pair(&o, &g, alpha, nalpha, gx[lvl], gw[lvl], orders[lvl], origin,
                                      edges, bins ? nbins + 1 : 0, pb);
                }
                double change = fabs(value - previous);
                double l1 = 0.;
                if (pb) {
                    for (int b = 0; b < nbins; ++b) l1 += fabs(pb[b] - pb2[b]);
                    memcpy(pb2, pb, sizeof(double) * nbins);
                }
                last_l1 = l1;
                charge_converged = m > 0 && change <= rel_tol * fabs(value) + abs_tol;
                if (charge_converged && l1 <= bin_tol * fabs(value) + abs_tol) {
                    converged = 1;
                    break;
                }
                previous = value;
            }
            /* charge must converge; bins that reach the last level are kept
             * and their last change is reported as their uncertainty */
            if (!converged && !charge_converged) failed = 1;
            if (bins_change) bins_change[om] += last_l1;
            charge[om] += value;
            if (rowb)
                for (int b = 0; b < nbins; ++b) rowb[b] += pb[b];
        }
        free(c);
        free(pb);
        free(pb2);
    }
    free(first);
    return failed;
}
