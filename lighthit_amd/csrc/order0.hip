// Stage (a): order-0 (direct) light to the first glass entry of every OM.
//
// The integral of PreparedBallistic (_surface_ballistic.py:20-353): photons
// of every element (straight chord, Cherenkov cone cos(theta) = 1/(beta n))
// that reach the OM sphere, weighted by acceptance, extinction over the entry
// distance and the detector spectral efficiency,
//
//   q0 = area/(pi R^2) sum_i int ds int dlambda S_i(lambda)
//        (1/2 pi) int dphi A(d.look) exp(-mu(lambda) d_entry).
//
// Formulation (cap coordinates): at an emission point s the OM is a cap of
// angular radius eps = asin(R/D) around the direction psi of its centre, and
// d lambda = |d lambda / d theta| d theta inside a medium panel. With
// theta = psi + eps sin(t) and phi = half sin(eta) every integrand is smooth,
// so fixed Gauss rules converge geometrically:
//   s:     6 nodes per piece of [0, L], split where a medium-panel edge is
//          tangent to the cap (s = z - rho cot(theta_k) -+ R / sin(theta_k));
//   theta: 8 nodes per medium panel inside the cap;
//   phi:   8 nodes on [0, half], both azimuth signs.
// The detector efficiency stays the verified cubic of its own (finer) panel.
//
// Launch shape: one workgroup per (OM, chunk of elements). Its threads test
// the chunk's (OM, element) pairs (the candidate test of interval_candidates,
// _ballistic_interval_numba.py:450-519, in FP64) and compact the survivors in
// LDS in a fixed order; then each wave integrates one pair at a time with its
// 64 lanes on the 8 x 8 (theta, phi) nodes. The per-(s, panel) set-up (FP64:
// the cap, the panel lines relative to it) is computed by one lane each and
// broadcast with shuffles; the node arithmetic is FP32 and formed from
// differences that do not cancel (haversines of the cap offset and azimuth).
// Per-lane sums are FP64 across pairs; waves and chunks combine in a fixed
// order, so the result does not depend on scheduling.
#include "lh_common.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxNs = 6;          // most s nodes per s piece
constexpr int kNt = 8;             // theta nodes per medium panel
constexpr int kNe = 8;             // azimuth nodes on [0, pi/2] (4 for small caps)
constexpr int kMaxMajor = 12;      // medium nodes inside the band (+ ends)
constexpr int kMaxPanels = 256;    // efficiency panels
constexpr int kMaxAlpha = 8;
constexpr double kC = 0.299792458;
constexpr int kChunk = 512;        // most elements per (OM, chunk) item
constexpr int kBoundWords = 12;
constexpr size_t kPersistentGroups = 2048;  // workgroups of the pair kernel
constexpr double kPi = 3.14159265358979323846;

// 8-point Gauss-Legendre rule on [-1, 1]
__constant__ float kGauss8x[8] = {-0.960289856497536f, -0.796666477413627f, -0.525532409916329f,
                                  -0.183434642495650f, 0.183434642495650f, 0.525532409916329f,
                                  0.796666477413627f, 0.960289856497536f};
__constant__ float kGauss4x[4] = {-0.861136311594053f, -0.339981043584856f, 0.339981043584856f,
                                  0.861136311594053f};
__constant__ float kGauss4w[4] = {0.347854845137454f, 0.652145154862546f, 0.652145154862546f,
                                  0.347854845137454f};
__constant__ float kGauss8w[8] = {0.101228536290376f, 0.222381034453374f, 0.313706645877887f,
                                  0.362683783378362f, 0.362683783378362f, 0.313706645877887f,
                                  0.222381034453374f, 0.101228536290376f};

// Gauss-Legendre rules on [-1, 1] of 2, 3, 4 and 6 points (the s rules),
// stored one after the other from kGaussSOffset[n].
__constant__ double kGaussSx[15] = {
    -0.577350269189626, 0.577350269189626,
    -0.774596669241483, 0.0, 0.774596669241483,
    -0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053,
    -0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197,
    0.661209386466264, 0.932469514203152};
__constant__ double kGaussSw[15] = {
    1.0, 1.0,
    0.555555555555556, 0.888888888888889, 0.555555555555556,
    0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454,
    0.171324492379170, 0.360761573048139, 0.467913934572691, 0.467913934572691,
    0.360761573048139, 0.171324492379170};
__device__ __forceinline__ int gauss_s_offset(int n) { return n == 2 ? 0 : n == 3 ? 2 : n == 4 ? 5 : 9; }

// Azimuth rule: eta = pi/4 (x + 1) on [0, pi/2] with the 8-point Gauss nodes;
// sin, cos, 1 - sin (cancellation-free) and weight pi/4 w of every node.
__constant__ float kEtaSin[8] = {3.1183217838e-02f, 1.5901983477e-01f, 3.6408107699e-01f, 5.9826084449e-01f, 8.0130141766e-01f, 9.3136725806e-01f, 9.8727538820e-01f, 9.9951368521e-01f};
__constant__ float kEtaCos[8] = {9.9951368521e-01f, 9.8727538820e-01f, 9.3136725806e-01f, 8.0130141766e-01f, 5.9826084449e-01f, 3.6408107699e-01f, 1.5901983477e-01f, 3.1183217838e-02f};
__constant__ float kEtaOneMinus[8] = {9.6881678216e-01f, 8.4098016523e-01f, 6.3591892301e-01f, 4.0173915551e-01f, 1.9869858234e-01f, 6.8632741945e-02f, 1.2724611798e-02f, 4.8631478841e-04f};
__constant__ float kEtaWeight[8] = {7.9504706486e-02f, 1.7465765603e-01f, 2.4638462352e-01f, 2.8485117736e-01f, 2.8485117736e-01f, 2.4638462352e-01f, 1.7465765603e-01f, 7.9504706486e-02f};
// The same with the 4-point Gauss nodes (caps far smaller than the azimuth
// scale of the acceptance and attenuation).
__constant__ float kEta4Sin[4] = {1.0884720006e-01f, 4.9547160085e-01f, 8.6862413779e-01f, 9.9405849277e-01f};
__constant__ float kEta4Cos[4] = {9.9405849277e-01f, 8.6862413779e-01f, 4.9547160085e-01f, 1.0884720006e-01f};
__constant__ float kEta4OneMinus[4] = {8.9115279994e-01f, 5.0452839915e-01f, 1.3137586221e-01f, 5.9415072342e-03f};
__constant__ float kEta4Weight[4] = {2.7320455650e-01f, 5.1219360690e-01f, 5.1219360690e-01f, 2.7320455650e-01f};

// One event's elements and its place in a launch: its chunks are the
// launch's chunks chunk_base .. chunk_base + nchunk - 1 and its (OM, chunk)
// items item_base + om * nchunk + chunk. An event without elements has no
// chunks; `out` (its charges) is null for padding entries of a table.
struct O0Event {
    const double* start;      // (nel, 3)
    const double* dir;        // (nel, 3)
    const double* length;
    const double* c0;
    const double* c2;
    const double* beta;
    const double* start_ns;   // (nel), bins only
    const double* end_ns;     // (nel), bins only
    double* out;              // (nom * width) result, or null
    long long chunk_base, item_base;
    int nel, chunk, nchunk;
};

constexpr int kEventsPerSet = 16;   // table entries written by one set-up launch

struct O0EventSet {
    O0Event e[kEventsPerSet];
};

struct O0Params {
    // OMs
    const double* positions;  // (nom, 3)
    const double* looks;      // (nom, 3)
    const double* areas;      // (nom)
    const double* radii;      // (nom)
    int nom;
    // events (a table in the workspace) and the launch's capacities
    const O0Event* events;
    int nevents;
    long long chunk_capacity, item_capacity;
    // optics panels (nm, phase, group, extinction at edges; cubic efficiency)
    const double* edges;
    const double* phase;
    const double* group;
    const double* ext;
    const double* eff;        // (npanel, effcols)
    int npanel, effcols;
    const int* major;         // panel-edge indices of the medium nodes in the band
    int nmajor;
    const double* alpha;
    int nalpha;
    double phase_min, phase_max;
    double* partial;          // (live items, width): one row per live item
    double* bounds;           // (nchunk, kBoundWords) bounding sphere / direction cone
    int* live;                // (items) live items (their launch index), in launch order
    int* live_flag;           // (item capacity + 1) 1 where the item may hold a pair
    int* live_slot;           // (item capacity + 1) exclusive scan of live_flag
    int* live_count;          // = live_slot + item capacity
    int* range;               // (live items, 2) first and last bin an item wrote
    // bins only (one event)
    const double* origins;    // (nom) time origins
    const double* bin_edges;  // (nbins + 1)
    int nbins;
    int* flags;               // [0]: glass intersected, [1]: unsupported geometry
    unsigned long long* node_count;   // optional: (theta, azimuth) nodes evaluated
    int s_rule;               // 0: s nodes per pair from its length; else that many
    int azimuth_rule;         // 0: azimuth nodes per pair from its distance; else that many
};

// Device counter of evaluated quadrature nodes (null: not counted); set by
// lh_order0_count_nodes and read by every later launch in the process.
static unsigned long long* g_node_count = nullptr;

// Rules of later launches (0: chosen per pair; s: 2, 3, 4, 6 and azimuth: 4, 8
// forced), set by lh_order0_set_rules.
static int g_s_rule = 0;
static int g_azimuth_rule = 0;

// Launch totals of a table (the last entry ends the launch's chunks and items).
__device__ __forceinline__ long long total_chunks(const O0Params& P) {
    const O0Event& last = P.events[P.nevents - 1];
    return last.chunk_base + last.nchunk;
}
__device__ __forceinline__ long long total_items(const O0Params& P) {
    const O0Event& last = P.events[P.nevents - 1];
    return last.item_base + (long long)P.nom * last.nchunk;
}

// The event of a launch chunk / item (the last entry starting at or before it).
__device__ __forceinline__ int event_of_chunk(const O0Params& P, long long c) {
    int e = 0;
    while (e + 1 < P.nevents && P.events[e + 1].chunk_base <= c) ++e;
    return e;
}
__device__ __forceinline__ int event_of_item(const O0Params& P, long long idx) {
    int e = 0;
    while (e + 1 < P.nevents && P.events[e + 1].item_base <= idx) ++e;
    return e;
}

// ---------------------------------------------------------- candidate test

// _cone_can_reach_sphere (_ballistic_interval_numba.py:226-252)
__device__ bool cone_can_reach_sphere(double z, double rho, double length, double beta,
                                      double radius, double phase_min, double phase_max) {
    if (z <= -radius || beta * phase_max <= 1.) return false;
    const double c_hi = 1. / (beta * phase_max);
    const double s_hi = sqrt(fmax(0., (1. - c_hi) * (1. + c_hi)));
    const double c_lo = fmin(1., 1. / (beta * phase_min));
    const double s_lo = sqrt(fmax(0., (1. - c_lo) * (1. + c_lo)));
    const double q_hi = rho * c_hi - z * s_hi;
    const double minimum = z >= 0. ? q_hi : fmin(q_hi, rho * c_lo - z * s_lo);
    const double rounding = 16 * 2.220446049250313e-16 * (fabs(z) + rho + length + radius);
    if (minimum > radius + rounding) return false;
    const double end_z = z - length;
    if (end_z >= 0. && rho * c_lo - end_z * s_lo < -radius - rounding) return false;
    return true;
}

// Candidate test of interval_candidates for a sphere (infinite reach) with the
// support test of _support (:152-186). Returns 1 for a pair to integrate.
__device__ int candidate(const O0Params& P, const O0Event& E, int om, int i, int* glass) {
    if (E.c0[i] == 0. || E.beta[i] * P.phase_max <= 1.) return 0;
    const double R = P.radii[om];
    const double dx = P.positions[3 * om] - E.start[3 * i];
    const double dy = P.positions[3 * om + 1] - E.start[3 * i + 1];
    const double dz = P.positions[3 * om + 2] - E.start[3 * i + 2];
    const double ux = E.dir[3 * i], uy = E.dir[3 * i + 1], uz = E.dir[3 * i + 2];
    const double z = dx * ux + dy * uy + dz * uz;
    if (z <= 0. && z <= -R) return 0;
    const double cx = dy * uz - dz * uy, cy = dz * ux - dx * uz, cz = dx * uy - dy * ux;
    const double rho = sqrt(cx * cx + cy * cy + cz * cz);
    const double L = E.length[i];
    const double along = fmax(fmax(-z, z - L), 0.);
    if (rho != 0.) {
        if (!cone_can_reach_sphere(z, rho, L, E.beta[i], R, P.phase_min, P.phase_max)) return 0;
        const double quarter = 0.5 * kPi;
        const double psi0 = atan2(rho, z), psi1 = atan2(rho, z - L);
        const double a0 = asin(fmin(1., R / hypot(z, rho)));
        const double a1 = asin(fmin(1., R / hypot(z - L, rho)));
        const double low = psi0 - a0, high = psi1 + a1;
        if (low >= quarter) return 0;
        const double b = E.beta[i];
        const double nmin = low <= 0. ? (1. + 1e-9) / b : fmax((1. + 1e-9) / b, 1. / (b * cos(low)));
        const double nmax = high >= quarter ? INFINITY : 1. / (b * cos(high));
        if (nmin > P.phase_max || nmax <= P.phase_min) return 0;
    }
    if (hypot(rho, along) <= R) *glass = 1;
    return 1;
}

// ------------------------------------------------------ node arithmetic

// Bonnet's recurrence l P_l = (2l - 1) x P_{l-1} - (l - 1) P_{l-2} with its
// two ratios tabulated (no division in the node loop).
__constant__ float kBonnetA[kMaxAlpha] = {0.f, 0.f, 3.f / 2.f, 5.f / 3.f, 7.f / 4.f,
                                          9.f / 5.f, 11.f / 6.f, 13.f / 7.f};
__constant__ float kBonnetB[kMaxAlpha] = {0.f, 0.f, 1.f / 2.f, 2.f / 3.f, 3.f / 4.f,
                                          4.f / 5.f, 5.f / 6.f, 6.f / 7.f};

__device__ __forceinline__ float acceptance(const float* a, int n, float x) {
    // sum (2l+1) alpha_l P_l(x) / (4 pi), clipped at zero (_ballistic_interval_numba.py:21-37)
    float result = a[0];
    if (n > 1) result += 3.f * a[1] * x;
    float previous = 1.f, current = x;
    for (int ell = 2; ell < n; ++ell) {
        const float value = kBonnetA[ell] * x * current - kBonnetB[ell] * previous;
        result += (2 * ell + 1) * a[ell] * value;
        previous = current;
        current = value;
    }
    return fmaxf(result, 0.f);
}

// 1 - sin(t) and 1 + sin(t) without cancellation near t = +-pi/2
__device__ __forceinline__ void one_minus_plus(float st, float ct, float& om, float& op) {
    if (st > 0.f) {
        op = 1.f + st;
        om = ct * ct / op;
    } else {
        om = 1.f - st;
        op = ct * ct / om;
    }
}

struct Slot {
    // one (s node, medium panel) of a pair; FP64 set-up rounded to FP32
    float weight;     // s weight * area / (pi R^2)
    float D, D2R2, eps, hav_eps, sinpsi;
    float psi_theta_a;  // psi - theta_a
    float theta_a, cos_theta_a;
    float ta, tb;
    float base;       // emission time minus the OM time origin [ns] (bins only)
    float DmR;        // D - R
    int piece;        // medium panel index, -1: empty
};

// Set-up of slot `slot` of a pair: s node slot / pieces of the s pieces,
// medium panel slot % pieces.
// Nodes of the s rule on every s piece of a pair. The charge takes 3 when
// the element is short against the glass radius (the cap then moves by at
// most a tenth of its own radius along the element, and 3 nodes agree with 6
// to the FP32 rounding of the node arithmetic), 6 otherwise; the bins keep 6
// (a bin edge crossed inside a piece leaves a kink in s). `rule` > 0 forces
// that many (2, 3, 4 or 6) for the convergence check.
template <bool kBins>
__device__ __forceinline__ int s_nodes(double L, double R, int rule) {
    if (rule > 0) return rule;
    return (!kBins && L <= 0.1 * R) ? 3 : kMaxNs;
}

// Azimuth nodes of a pair (per azimuth piece for the bins): 4 when the glass
// is at most 1/20 of the OM's distance from the element's line (the cap's
// azimuth half-width is then below ~0.08 rad, where the acceptance and
// attenuation are nearly linear and 4 Gauss nodes in eta agree with 8 to
// below 1e-7), 8 otherwise. `rule` > 0 forces 4 or 8.
__device__ __forceinline__ int azimuth_nodes(double rho, double R, int rule) {
    if (rule > 0) return rule;
    return R <= 0.05 * rho ? 4 : kNe;
}

__device__ Slot make_slot(const O0Params& P, int slot, int pieces, const double* scut, int nscut,
                          int ns, double z, double rho, double L, double R, double beta,
                          const double* theta_major, double norm, double t_start, double tau) {
    Slot out;
    out.piece = -1;
    const int snode = slot / pieces;
    const int piece = slot - snode * pieces;
    const int spiece = snode / ns;
    const int j = snode - spiece * ns + gauss_s_offset(ns);
    if (spiece >= nscut - 1) return out;
    const double sa = scut[spiece], sb = scut[spiece + 1];
    if (!(sb > sa)) return out;
    // ns-point Gauss-Legendre on [sa, sb]
    const double s = 0.5 * (sa + sb) + 0.5 * (sb - sa) * kGaussSx[j];
    const double along = z - s;
    const double D = sqrt(rho * rho + along * along);
    const double psi = atan2(rho, along);
    const double eps = asin(R / D);
    const double th0 = theta_major[piece], th1 = theta_major[piece + 1];
    const double lo = fmin(th0, th1), hi = fmax(th0, th1);
    const double xa = (lo - psi) / eps, xb = (hi - psi) / eps;
    if (xb <= -1. || xa >= 1.) return out;
    const double ta = xa <= -1. ? -0.5 * kPi : asin(xa);
    const double tb = xb >= 1. ? 0.5 * kPi : asin(xb);
    if (!(tb > ta)) return out;
    const double hs = sin(0.5 * eps);
    out.weight = (float)(0.5 * (sb - sa) * kGaussSw[j] * norm);
    out.D = (float)D;
    out.D2R2 = (float)((D - R) * (D + R));
    out.eps = (float)eps;
    out.hav_eps = (float)(hs * hs);
    out.sinpsi = (float)(rho / D);
    out.psi_theta_a = (float)(psi - th0);
    out.theta_a = (float)th0;
    out.cos_theta_a = (float)cos(th0);
    out.ta = (float)ta;
    out.tb = (float)tb;
    out.base = (float)(t_start + tau * s);
    out.DmR = (float)(D - R);
    out.piece = piece;
    return out;
}

// A non-empty slot queued for evaluation with the constants of its pair.
struct Task {
    Slot S;
    float beta, c0, c2, inv_length, axial, across, side;
    int ne;          // azimuth nodes of the pair
};

constexpr int kTaskQueue = 72;   // at most 7 left over + one group of 64 slots

struct PanelTables {
    // medium panels between consecutive major lines
    float lam_a[kMaxMajor], dlam_dn[kMaxMajor], n_a[kMaxMajor];
    float mu_a[kMaxMajor], dmu[kMaxMajor], span[kMaxMajor];
    float g_a[kMaxMajor], dg[kMaxMajor];
    // efficiency panels
    float edges[kMaxPanels + 1];
    float eff[kMaxPanels][4];
    float alpha[kMaxAlpha];
};

__device__ __forceinline__ int panel_of(const PanelTables& T, int npanel, float lam) {
    int lo = 0, hi = npanel;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (T.edges[mid] <= lam) lo = mid; else hi = mid;
    }
    return lo;
}

// sin(x) for the small angles of the cap (|x| < 0.3: odd Taylor series to
// x^11, relative error < 1e-9), the library sinf otherwise.
__device__ __forceinline__ float sin_small(float x) {
    if (fabsf(x) < 0.3f) {
        const float x2 = x * x;
        return x * (1.f + x2 * (-1.f / 6.f + x2 * (1.f / 120.f + x2 * (-1.f / 5040.f
                    + x2 * (1.f / 362880.f + x2 * (-1.f / 39916800.f))))));
    }
    return sinf(x);
}

// Value of one slot at the lane's theta node, summed over the 8 azimuth
// nodes (without the slot weight).
template <int NE>
__device__ float theta_node_value(const PanelTables& T, const Slot& S, int npanel, int nalpha,
                                  float beta, float c0, float c2, float inv_length, float axial,
                                  float across, float side, float xt, float wt) {
    const int k = S.piece;
    const float hw = 0.5f * (S.tb - S.ta);
    const float t = 0.5f * (S.ta + S.tb) + hw * xt;
    float st, ct;
    sincosf(t, &st, &ct);
    const float delta = S.eps * st;
    const float dth = S.psi_theta_a + delta;            // theta - theta_a
    const float theta = S.theta_a + dth;
    float sth, cth;
    sincosf(theta, &sth, &cth);
    // n - n_a = 2 sin((theta + theta_a)/2) sin((theta - theta_a)/2) / (beta cos(theta) cos(theta_a))
    const float dn = 2.f * sinf(S.theta_a + 0.5f * dth) * sin_small(0.5f * dth)
                     / (beta * cth * S.cos_theta_a);
    const float n = T.n_a[k] + dn;
    const float lam = T.lam_a[k] + dn * T.dlam_dn[k];
    const float qm = (lam - T.lam_a[k]) / T.span[k];
    const float mu = T.mu_a[k] + qm * T.dmu[k];
    const int e = panel_of(T, npanel, lam);
    const float qe = (lam - T.edges[e]) / (T.edges[e + 1] - T.edges[e]);
    const float eta = ((T.eff[e][3] * qe + T.eff[e][2]) * qe + T.eff[e][1]) * qe + T.eff[e][0];
    const float spectral = (c0 + c2 / (n * n)) * inv_length / (lam * lam) * eta;
    const float jac = beta * n * n * sth * fabsf(T.dlam_dn[k]);
    // azimuth half-width inside the cap: hav(half) = (hav eps - hav delta) / (sin theta sin psi)
    float om, op;
    one_minus_plus(st, ct, om, op);
    const float H = sin_small(0.5f * S.eps * om) * sin_small(0.5f * S.eps * op);
    const float sts = sth * S.sinpsi;
    const float hh = fminf(H / sts, 1.f);
    if (!(hh > 0.f)) return 0.f;
    const float half = 2.f * asinf(sqrtf(hh));
    const float hdel = sin_small(0.5f * delta);
    const float hav_d = hdel * hdel;
    const float rest = 1.f - S.hav_eps;
    float sum = 0.f;
#pragma unroll
    for (int ie = 0; ie < NE; ++ie) {
        const float se = NE == 4 ? kEta4Sin[ie] : kEtaSin[ie];
        const float ce = NE == 4 ? kEta4Cos[ie] : kEtaCos[ie];
        const float em = NE == 4 ? kEta4OneMinus[ie] : kEtaOneMinus[ie];
        const float ep = 1.f + se;
        const float phi = half * se;
        const float hphi = sin_small(0.5f * phi);
        const float hav_g = hav_d + sts * hphi * hphi;
        // hav(eps) - hav(gamma) = sin(theta) sin(psi) sin((half - phi)/2) sin((half + phi)/2)
        const float gap = sts * sin_small(0.5f * half * em) * sin_small(0.5f * half * ep);
        const float inside = fmaxf(gap * (rest - hav_g), 0.f);
        const float d = __fdividef(S.D2R2,
                                   S.D * (1.f - 2.f * hav_g) + 2.f * S.D * sqrtf(inside));
        // cos(phi) = 1 - 2 hav(phi), sin(phi) = 2 sin(phi/2) cos(phi/2)
        const float cp = 1.f - 2.f * hphi * hphi;
        const float sp = 2.f * hphi * sqrtf(fmaxf(0.f, 1.f - hphi * hphi));
        const float acc = acceptance(T.alpha, nalpha, cth * axial + sth * (cp * across + sp * side))
                        + acceptance(T.alpha, nalpha, cth * axial + sth * (cp * across - sp * side));
        sum += half * ce * (NE == 4 ? kEta4Weight[ie] : kEtaWeight[ie]) * acc * __expf(-mu * d);
    }
    return hw * wt * S.eps * ct * spectral * jac * sum;
}

// Quantities of a slot at one theta (t) that the bins need.
struct ThetaPoint {
    float st, ct, sth, cth, n, lam, qm, mu, ng_c, sts, half, hav_d;
    float d0;      // entry distance on the centre meridian (phi = 0)
    bool valid;
};

__device__ __forceinline__ ThetaPoint theta_point(const PanelTables& T, const Slot& S, float beta,
                                                  float t) {
    ThetaPoint p;
    const int k = S.piece;
    sincosf(t, &p.st, &p.ct);
    const float delta = S.eps * p.st;
    const float dth = S.psi_theta_a + delta;
    const float theta = S.theta_a + dth;
    sincosf(theta, &p.sth, &p.cth);
    const float dn = 2.f * sinf(S.theta_a + 0.5f * dth) * sin_small(0.5f * dth)
                     / (beta * p.cth * S.cos_theta_a);
    p.n = T.n_a[k] + dn;
    p.lam = T.lam_a[k] + dn * T.dlam_dn[k];
    p.qm = (p.lam - T.lam_a[k]) / T.span[k];
    p.mu = T.mu_a[k] + p.qm * T.dmu[k];
    p.ng_c = (T.g_a[k] + p.qm * T.dg[k]) * (float)(1.0 / kC);
    float om, op;
    one_minus_plus(p.st, p.ct, om, op);
    const float H = sin_small(0.5f * S.eps * om) * sin_small(0.5f * S.eps * op);
    p.sts = p.sth * S.sinpsi;
    const float hh = fminf(H / p.sts, 1.f);
    p.valid = hh > 0.f;
    p.half = p.valid ? 2.f * asinf(sqrtf(hh)) : 0.f;
    const float hdel = sin_small(0.5f * delta);
    p.hav_d = hdel * hdel;
    const float rest = 1.f - S.hav_eps;
    p.d0 = S.D2R2 / (S.D * (1.f - 2.f * p.hav_d)
                     + 2.f * S.D * sqrtf(fmaxf((S.hav_eps - p.hav_d) * (rest - p.hav_d), 0.f)));
    return p;
}

// Arrival time on the centre meridian of a slot at t.
__device__ __forceinline__ float meridian_time(const PanelTables& T, const Slot& S, float beta,
                                               float t) {
    const ThetaPoint p = theta_point(T, S, beta, t);
    return S.base + p.d0 * p.ng_c;
}

// One theta node at t with weight wt_t (dtheta/dt and the rule weight
// included): the azimuth chord [0, half] is split at every bin edge its
// arrival time crosses (the time grows with the azimuth); each piece is
// integrated with the 8-point rule and added to the bin it lies in.
template <int NE>
__device__ void node_bins(const PanelTables& T, const Slot& S, int npanel, int nalpha, float beta,
                          float c0, float c2, float inv_length, float axial, float across,
                          float side, float t, float wt_t, float R, const float* edges, int nbins,
                          double* wave_bins) {
    const ThetaPoint p = theta_point(T, S, beta, t);
    if (!p.valid) return;
    const int e = panel_of(T, npanel, p.lam);
    const float qe = (p.lam - T.edges[e]) / (T.edges[e + 1] - T.edges[e]);
    const float eta = ((T.eff[e][3] * qe + T.eff[e][2]) * qe + T.eff[e][1]) * qe + T.eff[e][0];
    const float spectral = (c0 + c2 / (p.n * p.n)) * inv_length / (p.lam * p.lam) * eta;
    const float jac = beta * p.n * p.n * p.sth * fabsf(T.dlam_dn[S.piece]);
    const float prefactor = S.weight * wt_t * S.eps * p.ct * spectral * jac;
    const float rest = 1.f - S.hav_eps;
    const float drim = S.D * (1.f - 2.f * S.hav_eps);
    const float t0 = S.base + p.d0 * p.ng_c, t1 = S.base + drim * p.ng_c;
    int lo = 0, hi = nbins + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edges[mid] <= t0) lo = mid + 1; else hi = mid;
    }
    int edge = lo;
    float eta_a = 0.f;
    for (;;) {
        float eta_b = 0.5f * 3.14159265358979f;
        const bool split = edge <= nbins && edges[edge] < t1;
        if (split) {
            const float dT = (edges[edge] - S.base) / p.ng_c;
            const float hg = (dT - S.DmR) * (S.DmR + 2.f * R - dT) / (4.f * S.D * dT);
            const float hp = (hg - p.hav_d) / p.sts;
            const float x = hp <= 0.f ? 0.f
                                      : fminf(2.f * asinf(fminf(sqrtf(hp), 1.f)) / p.half, 1.f);
            eta_b = fmaxf(asinf(x), eta_a);
        }
        const int bin = edge - 1;
        if (eta_b > eta_a && bin >= 0 && bin < nbins) {
            const float hwe = 0.5f * (eta_b - eta_a), mide = 0.5f * (eta_a + eta_b);
            float sum = 0.f;
#pragma unroll
            for (int ie = 0; ie < NE; ++ie) {
                float se, ce;
                sincosf(mide + hwe * (NE == 4 ? kGauss4x[ie] : kGauss8x[ie]), &se, &ce);
                const float em = ce * ce / (1.f + se), ep = 1.f + se;
                const float phi = p.half * se;
                const float hphi = sin_small(0.5f * phi);
                const float hav_g = p.hav_d + p.sts * hphi * hphi;
                const float gap = p.sts * sin_small(0.5f * p.half * em) * sin_small(0.5f * p.half * ep);
                const float inside = fmaxf(gap * (rest - hav_g), 0.f);
                const float d = __fdividef(S.D2R2,
                                           S.D * (1.f - 2.f * hav_g) + 2.f * S.D * sqrtf(inside));
                const float cp = 1.f - 2.f * hphi * hphi;
                const float sp = 2.f * hphi * sqrtf(fmaxf(0.f, 1.f - hphi * hphi));
                const float acc = acceptance(T.alpha, nalpha, p.cth * axial + p.sth * (cp * across + sp * side))
                                + acceptance(T.alpha, nalpha, p.cth * axial + p.sth * (cp * across - sp * side));
                sum += p.half * ce * hwe * (NE == 4 ? kGauss4w[ie] : kGauss8w[ie]) * acc
                       * __expf(-p.mu * d);
            }
            atomicAdd(&wave_bins[bin], (double)(prefactor * sum));
        }
        if (!split) break;
        eta_a = eta_b;
        ++edge;
    }
}

constexpr int kMaxCuts = 12;

// The Illinois method (regula falsi whose kept end has its value halved when
// that end is kept twice in a row) for the crossing of the centre meridian's
// arrival time with `Tedge` on [a, b], where the bracket keeps the sign change:
// every step stays inside the sub-interval; 8 steps reach the FP32 resolution
// of the times.
__device__ __forceinline__ float meridian_crossing(const PanelTables& T, const Slot& S, float beta,
                                                   float Tedge, float a, float b, float fa,
                                                   float fb) {
    float c = a;
    int kept = 0;
    for (int it = 0; it < 8; ++it) {
        c = (a * fb - b * fa) / (fb - fa);
        const float fc = meridian_time(T, S, beta, c) - Tedge;
        if (fc == 0.f) break;
        if ((fc < 0.f) == (fb < 0.f)) {
            b = c;
            fb = fc;
            if (kept == 1) fa *= 0.5f;
            kept = 1;
        } else {
            a = c;
            fa = fc;
            if (kept == 2) fb *= 0.5f;
            kept = 2;
        }
    }
    return c;
}

// Bins of one slot at the lane's theta node. The slot's theta range is cut
// where a bin edge crosses the centre meridian (sampled on 9 points, refined
// by the Illinois method) or the rim (there d_entry = D cos(eps) for every
// theta, so the crossing is where the group index takes one value: closed
// form); every piece gets the lane's node on a sine map, which absorbs the
// square-root behaviour of the bin weights at the cuts.
//
// The cuts depend on the slot only, so the 8 lanes of the slot's group (lanes
// `group` .. `group` + 7, one theta node each, `it` = lane - group) find them
// together: lane j evaluates sample j (every lane sample 8), the samples are
// exchanged, every lane lists the same crossings in the same (edge, interval)
// order, lane j refines crossings j, j + 8, ... and the refined cuts are
// exchanged. Every lane then holds the cuts a lone lane would find (the same
// expressions on the same values).
template <int NE>
__device__ void slot_bins(const PanelTables& T, const Slot& S, int npanel, int nalpha, float beta,
                          float c0, float c2, float inv_length, float axial, float across,
                          float side, float xt, float wt, float R, const float* edges, int nbins,
                          int it, int group, double* wave_bins) {
    const int k = S.piece;
    float cuts[kMaxCuts + 2];
    int ncut = 0;
    cuts[ncut++] = S.ta;
    float tv[9], tt[9];
    float lo_t = 3.4e38f, hi_t = -3.4e38f;
    const float own = meridian_time(T, S, beta, S.ta + (S.tb - S.ta) * (it * 0.125f));
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        tt[j] = S.ta + (S.tb - S.ta) * (j * 0.125f);
        tv[j] = j < 8 ? __shfl(own, group + j, 64) : meridian_time(T, S, beta, tt[j]);
        lo_t = fminf(lo_t, tv[j]);
        hi_t = fmaxf(hi_t, tv[j]);
    }
    // rim times at the panel's two group indices
    const float drim = S.D * (1.f - 2.f * S.hav_eps);
    const float ga = T.g_a[k] * (float)(1.0 / kC), gb = (T.g_a[k] + T.dg[k]) * (float)(1.0 / kC);
    lo_t = fminf(lo_t, S.base + drim * fminf(ga, gb));
    hi_t = fmaxf(hi_t, S.base + drim * fmaxf(ga, gb));
    int e = 0;
    {
        int a = 0, b = nbins + 1;
        while (a < b) {
            const int mid = (a + b) >> 1;
            if (edges[mid] <= lo_t) a = mid + 1; else b = mid;
        }
        e = a;
    }
    // meridian crossings to refine, packed as (edge << 8) | (cut slot << 4) |
    // interval j; rim cuts are set at once
    int task[kMaxCuts];
    int ntask = 0;
    for (; e <= nbins && edges[e] < hi_t && ncut < kMaxCuts; ++e) {
        const float Tedge = edges[e];
        for (int j = 0; j < 8 && ncut < kMaxCuts; ++j) {
            if ((tv[j] < Tedge) == (tv[j + 1] < Tedge)) continue;
            task[ntask++] = (e << 8) | (ncut << 4) | j;
            cuts[ncut++] = 0.f;
        }
        // rim crossing: n_g = (T - base) c / (D cos eps)
        if (T.dg[k] != 0.f && ncut < kMaxCuts) {
            const float q = ((Tedge - S.base) / drim * (float)kC - T.g_a[k]) / T.dg[k];
            if (q > 0.f && q < 1.f) {
                // theta of that wavelength, relative to theta_a: fixed point on
                // sin((th - th_a)/2) = (n - n_a) beta cos(th) cos(th_a) / (2 sin((th + th_a)/2))
                const float dn = q * T.span[k] / T.dlam_dn[k];
                float dth = 0.f;
                for (int iter = 0; iter < 3; ++iter) {
                    const float th = S.theta_a + dth;
                    dth = 2.f * asinf(dn * beta * cosf(th) * S.cos_theta_a
                                      / (2.f * sinf(S.theta_a + 0.5f * dth)));
                }
                const float x = (dth - S.psi_theta_a) / S.eps;
                if (x > -1.f && x < 1.f) {
                    const float tc = asinf(x);
                    if (tc > S.ta && tc < S.tb) cuts[ncut++] = tc;
                }
            }
        }
    }
    // lane `it` refines tasks it, it + 8 (at most kMaxCuts < 16 tasks)
    float refined[2] = {0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int q = it + 8 * r;
        if (q < ntask) {
            const int j = task[q] & 7;
            const float Tedge = edges[task[q] >> 8];
            refined[r] = meridian_crossing(T, S, beta, Tedge, tt[j], tt[j + 1], tv[j] - Tedge,
                                           tv[j + 1] - Tedge);
        }
    }
    for (int q = 0; q < ntask; ++q) {
        const float mine = q < 8 ? refined[0] : refined[1];
        cuts[(task[q] >> 4) & 15] = __shfl(mine, group + (q & 7), 64);
    }
    cuts[ncut++] = S.tb;
    for (int a = 1; a < ncut; ++a)  // insertion sort (few entries)
        for (int b = a; b > 0 && cuts[b - 1] > cuts[b]; --b) {
            const float tmp = cuts[b];
            cuts[b] = cuts[b - 1];
            cuts[b - 1] = tmp;
        }
    const float xm = 1.57079632679f * xt;
    float sxm, cxm;
    sincosf(xm, &sxm, &cxm);
    for (int piece = 0; piece + 1 < ncut; ++piece) {
        const float pa = cuts[piece], pb = cuts[piece + 1];
        if (!(pb > pa)) continue;
        const float hw = 0.5f * (pb - pa);
        node_bins<NE>(T, S, npanel, nalpha, beta, c0, c2, inv_length, axial, across, side,
                  0.5f * (pa + pb) + hw * sxm, hw * cxm * 1.57079632679f * wt, R, edges, nbins,
                  wave_bins);
    }
}

// Bounding sphere of the chunk's segments (centre, radius), mean direction
// and the cosine of the largest deviation from it, and the beta range.
__global__ __launch_bounds__(kThreads) void order0_chunk_bounds(O0Params P) {
    __shared__ double red[kThreads][4];
    const int tid = threadIdx.x;
    const long long c = blockIdx.x;
    if (c >= total_chunks(P)) return;
    const O0Event E = P.events[event_of_chunk(P, c)];
    const int first = (int)(c - E.chunk_base) * E.chunk;
    const int last = min(E.nel, first + E.chunk);
    double acc[4] = {0., 0., 0., 0.};
    double dir[3] = {0., 0., 0.};
    for (int i = first + tid; i < last; i += kThreads) {
        const double h = 0.5 * E.length[i];
        acc[0] += E.start[3 * i] + h * E.dir[3 * i];
        acc[1] += E.start[3 * i + 1] + h * E.dir[3 * i + 1];
        acc[2] += E.start[3 * i + 2] + h * E.dir[3 * i + 2];
        dir[0] += E.dir[3 * i];
        dir[1] += E.dir[3 * i + 1];
        dir[2] += E.dir[3 * i + 2];
    }
    // centre
    for (int k = 0; k < 3; ++k) red[tid][k] = acc[k];
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (tid < half)
            for (int k = 0; k < 3; ++k) red[tid][k] += red[tid + half][k];
        __syncthreads();
    }
    const double n = (double)(last - first);
    const double cx = red[0][0] / n, cy = red[0][1] / n, cz = red[0][2] / n;
    __syncthreads();
    for (int k = 0; k < 3; ++k) red[tid][k] = dir[k];
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (tid < half)
            for (int k = 0; k < 3; ++k) red[tid][k] += red[tid + half][k];
        __syncthreads();
    }
    double ux = red[0][0], uy = red[0][1], uz = red[0][2];
    const double un = sqrt(ux * ux + uy * uy + uz * uz);
    __syncthreads();
    // radius, smallest direction cosine, beta range
    double radius = 0., cosmin = 1., bmin = 1e300, bmax = 0.;
    if (un > 0.) {
        ux /= un;
        uy /= un;
        uz /= un;
    } else {
        cosmin = -1.;
    }
    for (int i = first + tid; i < last; i += kThreads) {
        const double L = E.length[i];
        for (int e = 0; e < 2; ++e) {
            const double x = E.start[3 * i] + e * L * E.dir[3 * i] - cx;
            const double y = E.start[3 * i + 1] + e * L * E.dir[3 * i + 1] - cy;
            const double z = E.start[3 * i + 2] + e * L * E.dir[3 * i + 2] - cz;
            radius = fmax(radius, sqrt(x * x + y * y + z * z));
        }
        cosmin = fmin(cosmin, E.dir[3 * i] * ux + E.dir[3 * i + 1] * uy + E.dir[3 * i + 2] * uz);
        bmin = fmin(bmin, E.beta[i]);
        bmax = fmax(bmax, E.beta[i]);
    }
    red[tid][0] = radius;
    red[tid][1] = cosmin;
    red[tid][2] = bmin;
    red[tid][3] = bmax;
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (tid < half) {
            red[tid][0] = fmax(red[tid][0], red[tid + half][0]);
            red[tid][1] = fmin(red[tid][1], red[tid + half][1]);
            red[tid][2] = fmin(red[tid][2], red[tid + half][2]);
            red[tid][3] = fmax(red[tid][3], red[tid + half][3]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        double* b = P.bounds + (size_t)c * kBoundWords;
        b[0] = cx; b[1] = cy; b[2] = cz; b[3] = red[0][0] * (1.0 + 1e-12) + 1e-9;
        b[4] = ux; b[5] = uy; b[6] = uz; b[7] = red[0][1];
        b[8] = red[0][2]; b[9] = red[0][3];
    }
}

// Conservative test that no element of a chunk can light the OM: every
// photon reaching the sphere leaves at an angle theta to its element with
// |theta - alpha| <= asin(r/D) + delta_u + asin(R/(D - r)), alpha the angle of
// the OM seen from the chunk centre against the mean direction.
__device__ bool chunk_culled(const O0Params& P, int chunk, int om) {
    const double* b = P.bounds + (size_t)chunk * kBoundWords;
    const double vx = P.positions[3 * om] - b[0], vy = P.positions[3 * om + 1] - b[1],
                 vz = P.positions[3 * om + 2] - b[2];
    const double D = sqrt(vx * vx + vy * vy + vz * vz);
    const double r = b[3], R = P.radii[om];
    if (!(D > r + R + 1e-6)) return false;
    const double dot = vx * b[4] + vy * b[5] + vz * b[6];
    const double cx = vy * b[6] - vz * b[5], cy = vz * b[4] - vx * b[6], cz = vx * b[5] - vy * b[4];
    const double alpha = atan2(sqrt(cx * cx + cy * cy + cz * cz), dot);
    const double delta = acos(fmin(1.0, fmax(-1.0, b[7])));
    const double slack = asin(fmin(1.0, r / D)) + delta + asin(fmin(1.0, R / (D - r))) + 1e-6;
    const double lo_bn = b[8] * P.phase_min, hi_bn = b[9] * P.phase_max;
    if (hi_bn <= 1.0) return true;
    const double theta_min = lo_bn > 1.0 ? acos(1.0 / lo_bn) : 0.0;
    const double theta_max = acos(1.0 / hi_bn);
    return alpha - slack > theta_max || alpha + slack < theta_min;
}

// Clears the flag words (status flags) of a call.
__global__ void order0_reset(int* flags) {
    if (threadIdx.x < 64) flags[threadIdx.x] = 0;
}

// Clears the flag words and writes `count` entries of the event table from
// `first` on.
__global__ void order0_set_events(int* flags, O0Event* table, int first, int count,
                                  O0EventSet set) {
    if (first == 0 && threadIdx.x < 64) flags[threadIdx.x] = 0;
    if (threadIdx.x < count) table[first + threadIdx.x] = set.e[threadIdx.x];
}

// Copies the flag word of a call to the caller's output.
__global__ void order0_copy_flags(const int* flags, int* out) {
    if (threadIdx.x == 0) out[0] = flags[0];
}

// One thread per (event, OM, chunk) item of the launch's capacity: flags the
// items that may hold a contributing pair (zero past the last item, for the
// scan).
__global__ void order0_cull(O0Params P) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx > P.item_capacity) return;
    if (idx >= total_items(P)) {
        P.live_flag[idx] = 0;
        return;
    }
    const O0Event& E = P.events[event_of_item(P, idx)];
    const long long local = idx - E.item_base;
    const int om = (int)(local / E.nchunk), chunk = (int)(local - (long long)om * E.nchunk);
    P.live_flag[idx] = chunk_culled(P, (int)(E.chunk_base + chunk), om) ? 0 : 1;
}

// Live items in launch order (event, OM, chunk) at their scanned slots: the
// list, and so every partial row, has the same order on every run.
__global__ void order0_compact(O0Params P) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P.item_capacity || !P.live_flag[idx]) return;
    P.live[P.live_slot[idx]] = (int)idx;
}

template <bool kBins>
__global__ __launch_bounds__(kThreads, 2) void order0_pairs(O0Params P) {
    extern __shared__ double dynamic_lds[];
    __shared__ PanelTables T;
    __shared__ int list[kChunk];
    __shared__ int wave_count[kWaves];
    __shared__ int total_count;
    __shared__ double wave_sum[kWaves];
    __shared__ double wave_theta[kWaves][kMaxMajor];
    __shared__ double wave_roots[kWaves][2 * kMaxMajor];
    __shared__ double wave_scut[kWaves][2 * kMaxMajor + 2];
    __shared__ int wave_nscut[kWaves];
    __shared__ Task wave_tasks[kWaves][kTaskQueue];
    __shared__ int glass;
    __shared__ int range_lo, range_hi;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // bins: one histogram per wave, then the bin edges (FP32)
    double* wave_bins = dynamic_lds + (size_t)wave * P.nbins;
    float* edges = reinterpret_cast<float*>(dynamic_lds + (size_t)kWaves * P.nbins);
    if (kBins) {
        for (int b = tid; b <= P.nbins; b += kThreads) edges[b] = (float)P.bin_edges[b];
    }

    // tables of this launch
    const int pieces = P.nmajor - 1;
    if (tid < pieces) {
        const int a = P.major[tid], b = P.major[tid + 1];
        T.lam_a[tid] = (float)P.edges[a];
        T.n_a[tid] = (float)P.phase[a];
        T.span[tid] = (float)(P.edges[b] - P.edges[a]);
        T.dlam_dn[tid] = (float)((P.edges[b] - P.edges[a]) / (P.phase[b] - P.phase[a]));
        T.mu_a[tid] = (float)P.ext[a];
        T.dmu[tid] = (float)(P.ext[b] - P.ext[a]);
        T.g_a[tid] = (float)P.group[a];
        T.dg[tid] = (float)(P.group[b] - P.group[a]);
    }
    for (int k = tid; k <= P.npanel; k += kThreads) T.edges[k] = (float)P.edges[k];
    for (int k = tid; k < P.npanel * 4; k += kThreads)
        T.eff[k / 4][k % 4] = (float)P.eff[(size_t)(k / 4) * P.effcols + (k % 4)];
    if (tid < P.nalpha) T.alpha[tid] = (float)(P.alpha[tid] / (4.0 * kPi));
    __syncthreads();

    // live (OM, chunk) items, strided over the persistent grid
    const int live = *P.live_count;
    for (int item = blockIdx.x; item < live; item += gridDim.x) {
    const long long idx = P.live[item];
    const O0Event E = P.events[event_of_item(P, idx)];
    const int om = (int)((idx - E.item_base) / E.nchunk);
    const int chunk = (int)(idx - E.item_base - (long long)om * E.nchunk);
    if (tid == 0) {
        total_count = 0;
        glass = 0;
    }
    if (kBins)
        for (int b = tid; b < kWaves * P.nbins; b += kThreads) dynamic_lds[b] = 0.0;
    __syncthreads();
    // candidate pairs of this chunk, compacted in element order
    const int first = chunk * E.chunk;
    const int last = min(E.nel, first + E.chunk);
    int local_glass = 0;
    for (int base = first; base < last; base += kThreads) {
        const int i = base + tid;
        const int pass = (i < last) ? candidate(P, E, om, i, &local_glass) : 0;
        const unsigned long long ballot = __ballot(pass);
        const int before = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wave_count[wave] = __popcll(ballot);
        __syncthreads();
        int offset = total_count;
        for (int w = 0; w < wave; ++w) offset += wave_count[w];
        if (pass) list[offset + before] = i;
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int w = 0; w < kWaves; ++w) sum += wave_count[w];
            total_count += sum;
        }
        __syncthreads();
    }
    if (local_glass) glass = 1;
    __syncthreads();
    if (tid == 0 && glass) atomicOr(&P.flags[0], 1);
    const int count = total_count;

    // integrate: one wave per pair; lane = (slot of a group of 8, theta node)
    const int it = lane & 7;             // theta node
    const int sub = lane >> 3;           // slot within the group of 8
    const float xt = kGauss8x[it], wt = kGauss8w[it];

    const double* pos = P.positions + 3 * om;
    const double* look = P.looks + 3 * om;
    const double R = P.radii[om];
    const double norm = P.areas[om] / (kPi * R * R) / (2.0 * kPi);
    double lane_sum = 0.0;
    unsigned long long lane_nodes = 0;
    // non-empty slots of the wave's pairs, evaluated 8 at a time (one per
    // group of 8 lanes, a theta node per lane) whatever pair they belong to
    Task* tasks = wave_tasks[wave];
    int queued = 0;
    auto wave_fence = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto evaluate = [&](int first, int n) {
        if (sub >= n) return;
        const Task& t = tasks[first + sub];
        lane_nodes += (unsigned long long)t.ne;
        if (kBins) {
            if (t.ne == 4)
                slot_bins<4>(T, t.S, P.npanel, P.nalpha, t.beta, t.c0, t.c2, t.inv_length,
                             t.axial, t.across, t.side, xt, wt, (float)R, edges, P.nbins, it,
                             lane - it, wave_bins);
            else
                slot_bins<kNe>(T, t.S, P.npanel, P.nalpha, t.beta, t.c0, t.c2, t.inv_length,
                               t.axial, t.across, t.side, xt, wt, (float)R, edges, P.nbins, it,
                               lane - it, wave_bins);
        } else {
            const float value = t.ne == 4
                ? theta_node_value<4>(T, t.S, P.npanel, P.nalpha, t.beta, t.c0, t.c2,
                                      t.inv_length, t.axial, t.across, t.side, xt, wt)
                : theta_node_value<kNe>(T, t.S, P.npanel, P.nalpha, t.beta, t.c0, t.c2,
                                        t.inv_length, t.axial, t.across, t.side, xt, wt);
            lane_sum += (double)(t.S.weight * value);
        }
    };
    // One loop over the groups of 64 slots of the wave's pairs: a group joins
    // the queue, then every full round of 8 queued slots is evaluated (all of
    // them once the last pair is done), so the evaluation has one call site.
    int entry = wave;           // the wave's next pair
    int group = 0, slots = 0;   // the current pair's next group and its slots
    double z = 0.0, rho = 0.0, L = 0.0, beta = 0.0, t_start = 0.0, tau = 0.0;
    float axial = 0.f, across = 0.f, side = 0.f, c0 = 0.f, c2 = 0.f, inv_length = 0.f;
    int ns = 0, ne = 0, nscut = 0;
    const double* scut = wave_scut[wave];
    const double* theta_major = wave_theta[wave];
    for (;;) {
        while (group >= slots && entry < count) {
            const int i = list[entry];
            entry += kWaves;
            const double ux = E.dir[3 * i], uy = E.dir[3 * i + 1], uz = E.dir[3 * i + 2];
            const double wx = pos[0] - E.start[3 * i], wy = pos[1] - E.start[3 * i + 1],
                         wz = pos[2] - E.start[3 * i + 2];
            z = wx * ux + wy * uy + wz * uz;
            const double tx = wx - z * ux, ty = wy - z * uy, tz = wz - z * uz;
            rho = sqrt(tx * tx + ty * ty + tz * tz);
            if (!(rho > R)) {
                if (lane == 0) atomicOr(&P.flags[0], 2);
                continue;
            }
            const double e1x = tx / rho, e1y = ty / rho, e1z = tz / rho;
            const double e2x = uy * e1z - uz * e1y, e2y = uz * e1x - ux * e1z,
                         e2z = ux * e1y - uy * e1x;
            axial = (float)(ux * look[0] + uy * look[1] + uz * look[2]);
            across = (float)(e1x * look[0] + e1y * look[1] + e1z * look[2]);
            side = (float)(e2x * look[0] + e2y * look[1] + e2z * look[2]);
            L = E.length[i];
            beta = E.beta[i];
            t_start = kBins ? E.start_ns[i] - P.origins[om] : 0.0;
            tau = kBins ? (E.end_ns[i] - E.start_ns[i]) / L : 0.0;
            // cone angles of the medium panel edges (one lane each) and the s
            // pieces: [0, L] cut where a panel edge is tangent to the cap
            if (lane < P.nmajor) {
                const double c = 1.0 / (beta * P.phase[P.major[lane]]);
                const double sn = sqrt(fmax(0.0, (1.0 - c) * (1.0 + c)));
                wave_theta[wave][lane] = atan2(sn, c);
                double r1 = -1.0, r2 = -1.0;
                if (sn > 0.0) {
                    const double centre = z - rho * c / sn, spread = R / sn;
                    r1 = centre - spread;
                    r2 = centre + spread;
                }
                wave_roots[wave][2 * lane] = r1;
                wave_roots[wave][2 * lane + 1] = r2;
            }
            wave_fence();
            if (lane == 0) {
                double* cut = wave_scut[wave];
                int n = 0;
                cut[n++] = 0.0;
                for (int m = 0; m < 2 * P.nmajor; ++m) {
                    const double r = wave_roots[wave][m];
                    if (r > 0.0 && r < L) cut[n++] = r;
                }
                cut[n++] = L;
                for (int a = 1; a < n; ++a)  // insertion sort (few entries)
                    for (int b = a; b > 0 && cut[b - 1] > cut[b]; --b) {
                        const double tmp = cut[b];
                        cut[b] = cut[b - 1];
                        cut[b - 1] = tmp;
                    }
                wave_nscut[wave] = n;
            }
            wave_fence();
            nscut = wave_nscut[wave];
            ns = s_nodes<kBins>(L, R, P.s_rule);
            ne = azimuth_nodes(rho, R, P.azimuth_rule);
            slots = (nscut - 1) * ns * pieces;
            group = 0;
            c0 = (float)E.c0[i];
            c2 = (float)E.c2[i];
            inv_length = (float)(1.0 / L);
        }
        const bool finished = group >= slots;   // no pair left
        if (!finished) {
            const Slot mine = (group + lane < slots)
                ? make_slot(P, group + lane, pieces, scut, nscut, ns, z, rho, L, R, beta,
                            theta_major, norm, t_start, tau)
                : Slot{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1};
            group += 64;
            // the group's non-empty slots (a cap meets few of the medium
            // panels) join the wave's queue in slot order
            const unsigned long long full = __ballot(mine.piece >= 0);
            if (mine.piece >= 0) {
                Task& t = tasks[queued + __popcll(full & ((1ULL << lane) - 1ULL))];
                t.S = mine;
                t.beta = (float)beta;
                t.c0 = c0;
                t.c2 = c2;
                t.inv_length = inv_length;
                t.axial = axial;
                t.across = across;
                t.side = side;
                t.ne = ne;
            }
            queued += __popcll(full);
        }
        wave_fence();
        int done = 0;
        for (; queued - done >= 8 || (finished && queued > done); done += 8)
            evaluate(done, min(8, queued - done));
        // the rest (fewer than 8) moves to the front of the queue
        const int left = max(queued - done, 0);
        if (done > 0 && left > 0) {
            Task keep;
            if (lane < left) keep = tasks[done + lane];
            wave_fence();
            if (lane < left) tasks[lane] = keep;
        }
        queued = left;
        wave_fence();
        if (finished) break;
    }
    if (P.node_count != nullptr) {   // one theta node per lane, ne azimuth nodes each
        for (int offset = 32; offset > 0; offset >>= 1) lane_nodes += __shfl_down(lane_nodes, offset);
        if (lane == 0 && lane_nodes) atomicAdd(P.node_count, lane_nodes);
    }
    if (kBins) {
        // waves in a fixed order into the first histogram; only the bins from
        // the first to the last nonzero one are written (the reduction reads
        // the item's range)
        if (tid == 0) {
            range_lo = P.nbins;
            range_hi = -1;
        }
        __syncthreads();
        int lo = P.nbins, hi = -1;
        for (int b = tid; b < P.nbins; b += kThreads) {
            double sum = 0.0;
            for (int w = 0; w < kWaves; ++w) sum += dynamic_lds[(size_t)w * P.nbins + b];
            dynamic_lds[b] = sum;
            if (sum != 0.0) {
                lo = min(lo, b);
                hi = max(hi, b);
            }
        }
        if (hi >= 0) {
            atomicMin(&range_lo, lo);
            atomicMax(&range_hi, hi);
        }
        __syncthreads();
        const int first_bin = range_lo, last_bin = range_hi;
        double* out = P.partial + (size_t)item * P.nbins;
        for (int b = first_bin + tid; b <= last_bin; b += kThreads) out[b] = dynamic_lds[b];
        if (tid == 0) {
            P.range[2 * item] = first_bin;
            P.range[2 * item + 1] = last_bin;
        }
        __syncthreads();
        continue;
    }
    // fixed-order reductions: lanes, then waves
    for (int offset = 32; offset > 0; offset >>= 1) lane_sum += __shfl_down(lane_sum, offset);
    if (lane == 0) wave_sum[wave] = lane_sum;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
        for (int w = 0; w < kWaves; ++w) sum += wave_sum[w];
        P.partial[item] = sum;
    }
    __syncthreads();
    }
}

// out[om][b] of every event = sum over the OM's live items, in chunk order,
// of their partial rows (width = bins or 1; a bin item adds only inside its
// range). An event without elements gets zeros; padding entries are skipped.
__global__ void order0_reduce(O0Params P, int width) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per_event = (long long)P.nom * width;
    if (idx >= per_event * P.nevents) return;
    const O0Event& E = P.events[idx / per_event];
    if (E.out == nullptr) return;
    const long long rest = idx % per_event;
    const int om = (int)(rest / width), b = (int)(rest - (long long)om * width);
    const long long at = E.item_base + (long long)om * E.nchunk;
    const int first = P.live_slot[at], last = P.live_slot[at + E.nchunk];
    double sum = 0.0;
    if (width == 1) {
        for (int item = first; item < last; ++item) sum += P.partial[item];
    } else {
        for (int item = first; item < last; ++item)
            if (b >= P.range[2 * item] && b <= P.range[2 * item + 1])
                sum += P.partial[(size_t)item * width + b];
    }
    E.out[rest] = sum;
}

}  // namespace

namespace {

// Elements per (OM, chunk) item, from the element count alone (so an OM's
// result never depends on which other OMs or events share the call): small
// sources are cut finer, so that few OMs near a short shower still give the
// chip several hundred workgroups and the pairs of the nearest OM spread over
// many waves. The bins launch (only the active OMs, each pair costlier) cuts
// small sources finer still.
int chunk_of(int n_elements, bool bins) {
    const int chunk = n_elements >= 32768 ? kChunk : n_elements >= 8192 ? 256
                    : n_elements >= 2048 ? 128 : n_elements >= 512 ? 64 : 32;
    return (bins && n_elements < 65536) ? std::max(16, chunk / 2) : chunk;
}

long long chunks_of(int n_elements, bool bins) {
    const int chunk = chunk_of(n_elements, bins);
    return n_elements > 0 ? (n_elements + chunk - 1) / chunk : 0;
}

size_t align256(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// Workspace of one launch: flag words | event table | chunk bounds | partial
// rows | live list | live flags | live slots | bin ranges | scan scratch.
struct O0Layout {
    long long chunks, items;
    size_t table, bounds, partial, live, flag, slot, range, scan, scan_bytes, total;
};

O0Layout layout_of(int n_oms, int max_events, long long max_chunks, int width, bool bins) {
    O0Layout L{};
    L.chunks = std::max<long long>(max_chunks, 0);
    L.items = (long long)(n_oms > 0 ? n_oms : 0) * L.chunks;
    const size_t w = (size_t)(width > 0 ? width : 0);
    const size_t items = (size_t)L.items;
    L.table = 256;
    L.bounds = align256(L.table + sizeof(O0Event) * (size_t)std::max(max_events, 1));
    L.partial = align256(L.bounds + sizeof(double) * (size_t)L.chunks * kBoundWords);
    L.live = align256(L.partial + sizeof(double) * items * w);
    L.flag = align256(L.live + sizeof(int) * items);
    L.slot = align256(L.flag + sizeof(int) * (items + 1));
    L.range = align256(L.slot + sizeof(int) * (items + 1));
    L.scan = align256(L.range + (bins ? 2 * sizeof(int) * items : 0));
    size_t scan_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int*)nullptr, (int*)nullptr,
                                         (int)(items + 1)) != hipSuccess)
        scan_bytes = 0;
    L.scan_bytes = std::max<size_t>(scan_bytes, 1);
    L.total = align256(L.scan + L.scan_bytes);
    return L;
}

int check_descriptors(const lh_order0_optics* optics, const lh_order0_oms* oms, const char* name) {
    LH_REQUIRE(optics && oms, "%s: null descriptor", name);
    LH_REQUIRE(oms->count >= 0, "%s: negative count", name);
    LH_REQUIRE(oms->count <= 65535, "%s: at most 65535 OMs per launch", name);
    LH_REQUIRE(optics->panels >= 1 && optics->panels <= kMaxPanels,
               "%s: 1..%d optical panels supported", name, kMaxPanels);
    LH_REQUIRE(optics->major_count >= 2 && optics->major_count <= kMaxMajor,
               "%s: 2..%d medium nodes in the band supported", name, kMaxMajor);
    LH_REQUIRE(optics->acceptance_count >= 1 && optics->acceptance_count <= kMaxAlpha,
               "%s: 1..%d acceptance coefficients supported", name, kMaxAlpha);
    LH_REQUIRE(optics->efficiency_columns == 4, "%s: cubic efficiency panels (4 columns) required",
               name);
    return LH_OK;
}

int check_descriptors(const lh_order0_optics* optics, const lh_order0_oms* oms,
                      const lh_order0_elements* elements, const char* name) {
    LH_REQUIRE(elements, "%s: null descriptor", name);
    LH_REQUIRE(elements->count >= 0, "%s: negative count", name);
    return check_descriptors(optics, oms, name);
}

// Table entries of `count` events (chunks and items laid out in event order),
// padded to `max_events` with entries that own nothing.
int event_table(const lh_order0_elements* elements, double* const* outs, int count,
                int max_events, int n_oms, bool bins, long long max_chunks, O0Event* table,
                const char* name) {
    long long chunks = 0;
    for (int k = 0; k < max_events; ++k) {
        O0Event E{};
        if (k < count) {
            const lh_order0_elements& el = elements[k];
            LH_REQUIRE(el.count >= 0, "%s: negative element count in event %d", name, k);
            LH_REQUIRE(el.count == 0 || (el.start_m && el.direction && el.length_m
                                         && el.coefficient0 && el.coefficient2 && el.beta),
                       "%s: null element array in event %d", name, k);
            E.start = el.start_m;
            E.dir = el.direction;
            E.length = el.length_m;
            E.c0 = el.coefficient0;
            E.c2 = el.coefficient2;
            E.beta = el.beta;
            E.start_ns = el.start_ns;
            E.end_ns = el.end_ns;
            E.out = outs[k];
            LH_REQUIRE(E.out != nullptr || n_oms == 0, "%s: null output of event %d", name, k);
            E.nel = el.count;
            E.chunk = chunk_of(el.count, bins);
            E.nchunk = (int)chunks_of(el.count, bins);
        }
        E.chunk_base = chunks;
        E.item_base = chunks * n_oms;
        chunks += E.nchunk;
        table[k] = E;
    }
    LH_REQUIRE(chunks <= max_chunks, "%s: %lld element chunks, above the launch's %lld", name,
               chunks, max_chunks);
    return LH_OK;
}

// Writes the table into the workspace (and clears the flag words) with
// kernels, never a host copy or a memset.
int upload_table(const O0Event* table, int count, void* workspace, hipStream_t s,
                 const char* name) {
    O0Event* dst = reinterpret_cast<O0Event*>(static_cast<char*>(workspace) + 256);
    for (int first = 0; first < std::max(count, 1); first += kEventsPerSet) {
        O0EventSet set{};
        const int n = std::min(kEventsPerSet, count - first);
        for (int k = 0; k < n; ++k) set.e[k] = table[first + k];
        hipLaunchKernelGGL(order0_set_events, dim3(1), dim3(64), 0, s,
                           static_cast<int*>(workspace), dst, first, std::max(n, 0), set);
        const int status = lh::check_launch(name);
        if (status != LH_OK) return status;
    }
    return LH_OK;
}

O0Params make_params(const lh_order0_optics* optics, const lh_order0_oms* oms, void* workspace,
                     int max_events, long long max_chunks, int width, bool bins) {
    O0Params P{};
    P.positions = oms->positions;
    P.looks = oms->looks;
    P.areas = oms->areas;
    P.radii = oms->radii;
    P.nom = oms->count;
    P.edges = optics->panel_edges;
    P.phase = optics->phase_nodes;
    P.group = optics->group_nodes;
    P.ext = optics->extinction_nodes;
    P.eff = optics->efficiency;
    P.npanel = optics->panels;
    P.effcols = optics->efficiency_columns;
    P.major = optics->major;
    P.nmajor = optics->major_count;
    P.alpha = optics->alpha;
    P.nalpha = optics->acceptance_count;
    P.phase_min = optics->phase_min;
    P.phase_max = optics->phase_max;
    const O0Layout L = layout_of(oms->count, max_events, max_chunks, width, bins);
    char* base = static_cast<char*>(workspace);
    P.events = reinterpret_cast<const O0Event*>(base + L.table);
    P.nevents = std::max(max_events, 1);
    P.chunk_capacity = L.chunks;
    P.item_capacity = L.items;
    P.flags = static_cast<int*>(workspace);
    P.node_count = g_node_count;
    P.s_rule = g_s_rule;
    P.azimuth_rule = g_azimuth_rule;
    P.bounds = reinterpret_cast<double*>(base + L.bounds);
    P.partial = reinterpret_cast<double*>(base + L.partial);
    P.live = reinterpret_cast<int*>(base + L.live);
    P.live_flag = reinterpret_cast<int*>(base + L.flag);
    P.live_slot = reinterpret_cast<int*>(base + L.slot);
    P.live_count = P.live_slot + L.items;
    P.range = bins ? reinterpret_cast<int*>(base + L.range) : nullptr;
    return P;
}

// The launch chain after the table is in place. Every grid is sized by the
// launch's capacities (never by the events of one call), and every size the
// kernels need is read from the table on the device: the chain can be
// captured once per capacity and replayed for any events that fit.
template <bool kBins>
int run(O0Params P, int width, void* workspace, hipStream_t s, size_t lds, const char* name) {
    const O0Layout L = layout_of(P.nom, P.nevents, P.chunk_capacity, width, kBins);
    if (L.chunks > 0) {
        hipLaunchKernelGGL(order0_chunk_bounds, dim3((unsigned)L.chunks), dim3(kThreads), 0, s, P);
        int status = lh::check_launch(name);
        if (status != LH_OK) return status;
    }
    hipLaunchKernelGGL(order0_cull, dim3((unsigned)((L.items + 1 + 255) / 256)), dim3(256), 0, s,
                       P);
    int status = lh::check_launch(name);
    if (status != LH_OK) return status;
    size_t scan_bytes = L.scan_bytes;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(static_cast<char*>(workspace) + L.scan,
                                                 scan_bytes, P.live_flag, P.live_slot,
                                                 (int)(L.items + 1), s));
    if (L.items > 0) {
        hipLaunchKernelGGL(order0_compact, dim3((unsigned)((L.items + 255) / 256)), dim3(256), 0,
                           s, P);
        status = lh::check_launch(name);
        if (status != LH_OK) return status;
        const unsigned grid = (unsigned)std::min<long long>(L.items, kPersistentGroups);
        lh::KernelMark mark(kBins ? "order0_pairs<true>" : "order0_pairs<false>", s);
        hipLaunchKernelGGL((order0_pairs<kBins>), dim3(grid), dim3(kThreads), lds, s, P);
    }
    status = lh::check_launch(name);
    if (status != LH_OK) return status;
    const long long total = (long long)P.nom * width * P.nevents;
    if (total > 0) {
        hipLaunchKernelGGL(order0_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           P, width);
        status = lh::check_launch(name);
    }
    return status;
}

// One event: its table entry, then the chain at its own capacity.
template <bool kBins>
int run_one(const lh_order0_optics* optics, const lh_order0_oms* oms,
            const lh_order0_elements* elements, double* out, int width, void* workspace,
            hipStream_t s, size_t lds, const double* origins, const double* bin_edges,
            const char* name) {
    O0Event table[1];
    const long long chunks = chunks_of(elements->count, kBins);
    int status = event_table(elements, &out, 1, 1, oms->count, kBins, chunks, table, name);
    if (status != LH_OK) return status;
    status = upload_table(table, 1, workspace, s, name);
    if (status != LH_OK) return status;
    O0Params P = make_params(optics, oms, workspace, 1, chunks, width, kBins);
    P.origins = origins;
    P.bin_edges = bin_edges;
    P.nbins = kBins ? width : 0;
    return run<kBins>(P, width, workspace, s, lds, name);
}

}  // namespace

extern "C" int lh_order0_count_nodes(unsigned long long* counter) {
    g_node_count = counter;
    return LH_OK;
}

extern "C" int lh_order0_set_rules(int s_nodes, int azimuth_nodes) {
    LH_REQUIRE(s_nodes == 0 || s_nodes == 2 || s_nodes == 3 || s_nodes == 4 || s_nodes == 6,
               "lh_order0_set_rules: s nodes 0 (per pair), 2, 3, 4 or 6 (got %d)", s_nodes);
    LH_REQUIRE(azimuth_nodes == 0 || azimuth_nodes == 4 || azimuth_nodes == 8,
               "lh_order0_set_rules: azimuth nodes 0 (per pair), 4 or 8 (got %d)", azimuth_nodes);
    g_s_rule = s_nodes;
    g_azimuth_rule = azimuth_nodes;
    return LH_OK;
}

extern "C" size_t lh_order0_bins_workspace_bytes(int n_oms, int n_elements, int n_bins) {
    return layout_of(n_oms, 1, chunks_of(n_elements, true), n_bins, true).total;
}

extern "C" size_t lh_order0_workspace_bytes(int n_oms, int n_elements) {
    return layout_of(n_oms, 1, chunks_of(n_elements, false), 1, false).total;
}

extern "C" int lh_order0_charge(const lh_order0_optics* optics, const lh_order0_oms* oms,
                                const lh_order0_elements* elements, double* charge,
                                void* workspace, size_t workspace_bytes, void* stream) {
    int status = check_descriptors(optics, oms, elements, "lh_order0_charge");
    if (status != LH_OK) return status;
    LH_REQUIRE(workspace_bytes >= lh_order0_workspace_bytes(oms->count, elements->count),
               "lh_order0_charge: workspace of %zu bytes, %zu needed", workspace_bytes,
               lh_order0_workspace_bytes(oms->count, elements->count));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (oms->count == 0) {
        hipLaunchKernelGGL(order0_reset, dim3(1), dim3(64), 0, s, static_cast<int*>(workspace));
        return lh::check_launch("lh_order0_charge (reset)");
    }
    return run_one<false>(optics, oms, elements, charge, 1, workspace, s, 0, nullptr, nullptr,
                          "lh_order0_charge");
}

extern "C" size_t lh_order0_batch_workspace_bytes(int n_oms, int max_events, long long max_chunks) {
    return layout_of(n_oms, max_events, max_chunks, 1, false).total;
}

extern "C" long long lh_order0_batch_chunks(const int* element_counts, int n_events) {
    long long chunks = 0;
    for (int k = 0; k < n_events; ++k) chunks += chunks_of(element_counts[k], false);
    return chunks;
}

extern "C" int lh_order0_batch_set(const lh_order0_oms* oms, const lh_order0_elements* events,
                                   double* const* charges, int n_events, int max_events,
                                   long long max_chunks, void* workspace, size_t workspace_bytes,
                                   void* stream) {
    const char* name = "lh_order0_batch_set";
    LH_REQUIRE(oms && events && charges, "%s: null argument", name);
    LH_REQUIRE(n_events >= 0 && n_events <= max_events && max_events >= 1,
               "%s: %d events for a launch of %d", name, n_events, max_events);
    LH_REQUIRE(max_events <= 4096, "%s: at most 4096 events per launch", name);
    LH_REQUIRE(oms->count >= 0 && oms->count <= 65535, "%s: 0..65535 OMs", name);
    const size_t need = lh_order0_batch_workspace_bytes(oms->count, max_events, max_chunks);
    LH_REQUIRE(workspace_bytes >= need, "%s: workspace of %zu bytes, %zu needed", name,
               workspace_bytes, need);
    O0Event* table = new O0Event[max_events];
    int status = event_table(events, charges, n_events, max_events, oms->count, false, max_chunks,
                             table, name);
    if (status == LH_OK)
        status = upload_table(table, max_events, workspace, static_cast<hipStream_t>(stream), name);
    delete[] table;
    return status;
}

extern "C" int lh_order0_charge_batch(const lh_order0_optics* optics, const lh_order0_oms* oms,
                                      int max_events, long long max_chunks, void* workspace,
                                      size_t workspace_bytes, void* stream) {
    const char* name = "lh_order0_charge_batch";
    int status = check_descriptors(optics, oms, name);
    if (status != LH_OK) return status;
    LH_REQUIRE(max_events >= 1 && max_events <= 4096 && max_chunks >= 0,
               "%s: 1..4096 events and a nonnegative chunk capacity", name);
    const size_t need = lh_order0_batch_workspace_bytes(oms->count, max_events, max_chunks);
    LH_REQUIRE(workspace_bytes >= need, "%s: workspace of %zu bytes, %zu needed", name,
               workspace_bytes, need);
    LH_REQUIRE((long long)oms->count * max_chunks < (1LL << 31),
               "%s: %lld (OM, chunk) items, above 2^31", name,
               (long long)oms->count * max_chunks);
    if (oms->count == 0) return LH_OK;
    O0Params P = make_params(optics, oms, workspace, max_events, max_chunks, 1, false);
    return run<false>(P, 1, workspace, static_cast<hipStream_t>(stream), 0, name);
}

extern "C" int lh_order0_bins(const lh_order0_optics* optics, const lh_order0_oms* oms,
                              const lh_order0_elements* elements, const double* origins,
                              const double* bin_edges, int n_bins, double* bins, void* workspace,
                              size_t workspace_bytes, void* stream) {
    int status = check_descriptors(optics, oms, elements, "lh_order0_bins");
    if (status != LH_OK) return status;
    LH_REQUIRE(n_bins >= 1 && n_bins <= 2048, "lh_order0_bins: 1..2048 bins supported (got %d)",
               n_bins);
    LH_REQUIRE(workspace_bytes >= lh_order0_bins_workspace_bytes(oms->count, elements->count, n_bins),
               "lh_order0_bins: workspace of %zu bytes, %zu needed", workspace_bytes,
               lh_order0_bins_workspace_bytes(oms->count, elements->count, n_bins));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (oms->count == 0) {
        hipLaunchKernelGGL(order0_reset, dim3(1), dim3(64), 0, s, static_cast<int*>(workspace));
        return lh::check_launch("lh_order0_bins (reset)");
    }
    LH_REQUIRE(origins && bin_edges && bins, "lh_order0_bins: null device pointer");
    const size_t lds = sizeof(double) * kWaves * (size_t)n_bins + sizeof(float) * (n_bins + 1);
    return run_one<true>(optics, oms, elements, bins, n_bins, workspace, s, lds, origins, bin_edges,
                         "lh_order0_bins");
}

// ---------------------------------------------------------------------------
// No-allocation charge entry point: the per-call workspace of lh_order0_charge
// lives in a fixed device array of this file. Every GPU holds its own copy of
// the array (64 MiB) from the library's first use on it until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device.

namespace {

constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(void** out) {
    int device = -1;
    LH_HIP_CALL(hipGetDevice(&device));
    LH_REQUIRE(device >= 0 && device < kMaxDevices,
               "lh_order0_charge_noalloc: device %d outside 0..%d", device, kMaxDevices - 1);
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (g_scratch_ptr[device] == nullptr) {
        void* ptr = nullptr;
        LH_HIP_CALL(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        LH_REQUIRE(ptr != nullptr, "lh_order0_charge_noalloc: no device address for scratch");
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return LH_OK;
}

}  // namespace

extern "C" size_t lh_order0_charge_noalloc_max_bytes(void) { return kNoallocScratchBytes; }

extern "C" int lh_order0_charge_noalloc(const lh_order0_optics* optics, const lh_order0_oms* oms,
                                        const lh_order0_elements* elements, double* charge,
                                        int* flags, void* stream) {
    int status = check_descriptors(optics, oms, elements, "lh_order0_charge_noalloc");
    if (status != LH_OK) return status;
    const size_t need = lh_order0_workspace_bytes(oms->count, elements->count);
    LH_REQUIRE(need <= kNoallocScratchBytes,
               "lh_order0_charge_noalloc: %d OMs x %d elements need %zu bytes of scratch, "
               "above its fixed %zu; call lh_order0_charge with a workspace instead",
               oms->count, elements->count, need, kNoallocScratchBytes);
    void* ws = nullptr;
    status = scratch_address(&ws);
    if (status != LH_OK) return status;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the flag words start at zero on every call (a kernel, not a memset)
    if (oms->count == 0) {
        hipLaunchKernelGGL(order0_reset, dim3(1), dim3(64), 0, s, static_cast<int*>(ws));
        status = lh::check_launch("lh_order0_charge_noalloc (reset)");
    } else {
        status = run_one<false>(optics, oms, elements, charge, 1, ws, s, 0, nullptr, nullptr,
                                "lh_order0_charge_noalloc");
    }
    if (status != LH_OK) return status;
    if (flags != nullptr) {
        hipLaunchKernelGGL(order0_copy_flags, dim3(1), dim3(64), 0, s,
                           static_cast<const int*>(ws), flags);
        status = lh::check_launch("lh_order0_charge_noalloc (flags)");
    }
    return status;
}
