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

#include <cmath>

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kNs = 6;             // s nodes per s piece
constexpr int kNt = 8;             // theta nodes per medium panel
constexpr int kNe = 8;             // azimuth nodes on [0, pi/2]
constexpr int kMaxMajor = 12;      // medium nodes inside the band (+ ends)
constexpr int kMaxPanels = 256;    // efficiency panels
constexpr int kMaxAlpha = 8;
constexpr double kC = 0.299792458;
constexpr double kPi = 3.14159265358979323846;

struct O0Params {
    // OMs
    const double* positions;  // (nom, 3)
    const double* looks;      // (nom, 3)
    const double* areas;      // (nom)
    const double* radii;      // (nom)
    int nom;
    // elements
    const double* start;      // (nel, 3)
    const double* dir;        // (nel, 3)
    const double* length;
    const double* c0;
    const double* c2;
    const double* beta;
    int nel;
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
    int chunk, nchunk;
    double* partial;          // (nom, nchunk)
    int* flags;               // [0]: glass intersected, [1]: unsupported geometry
};

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
__device__ int candidate(const O0Params& P, int om, int i, int* glass) {
    if (P.c0[i] == 0. || P.beta[i] * P.phase_max <= 1.) return 0;
    const double R = P.radii[om];
    const double dx = P.positions[3 * om] - P.start[3 * i];
    const double dy = P.positions[3 * om + 1] - P.start[3 * i + 1];
    const double dz = P.positions[3 * om + 2] - P.start[3 * i + 2];
    const double ux = P.dir[3 * i], uy = P.dir[3 * i + 1], uz = P.dir[3 * i + 2];
    const double z = dx * ux + dy * uy + dz * uz;
    if (z <= 0. && z <= -R) return 0;
    const double cx = dy * uz - dz * uy, cy = dz * ux - dx * uz, cz = dx * uy - dy * ux;
    const double rho = sqrt(cx * cx + cy * cy + cz * cz);
    const double L = P.length[i];
    const double along = fmax(fmax(-z, z - L), 0.);
    if (rho != 0.) {
        if (!cone_can_reach_sphere(z, rho, L, P.beta[i], R, P.phase_min, P.phase_max)) return 0;
        const double quarter = 0.5 * kPi;
        const double psi0 = atan2(rho, z), psi1 = atan2(rho, z - L);
        const double a0 = asin(fmin(1., R / hypot(z, rho)));
        const double a1 = asin(fmin(1., R / hypot(z - L, rho)));
        const double low = psi0 - a0, high = psi1 + a1;
        if (low >= quarter) return 0;
        const double b = P.beta[i];
        const double nmin = low <= 0. ? (1. + 1e-9) / b : fmax((1. + 1e-9) / b, 1. / (b * cos(low)));
        const double nmax = high >= quarter ? INFINITY : 1. / (b * cos(high));
        if (nmin > P.phase_max || nmax <= P.phase_min) return 0;
    }
    if (hypot(rho, along) <= R) *glass = 1;
    return 1;
}

// ------------------------------------------------------ node arithmetic

__device__ __forceinline__ float acceptance(const float* a, int n, float x) {
    // sum (2l+1) alpha_l P_l(x) / (4 pi), clipped at zero (_ballistic_interval_numba.py:21-37)
    float result = a[0];
    if (n > 1) result += 3.f * a[1] * x;
    float previous = 1.f, current = x;
    for (int ell = 2; ell < n; ++ell) {
        const float value = ((2 * ell - 1) * x * current - (ell - 1) * previous) / ell;
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
    int piece;        // medium panel index, -1: empty
};

// Set-up of slot `slot` of a pair: s node slot / pieces of the s pieces,
// medium panel slot % pieces.
__device__ Slot make_slot(const O0Params& P, int slot, int pieces, const double* scut, int nscut,
                          double z, double rho, double L, double R, double beta,
                          const double* theta_major, double norm) {
    Slot out;
    out.piece = -1;
    const int snode = slot / pieces;
    const int piece = slot - snode * pieces;
    const int spiece = snode / kNs;
    const int j = snode - spiece * kNs;
    if (spiece >= nscut - 1) return out;
    const double sa = scut[spiece], sb = scut[spiece + 1];
    if (!(sb > sa)) return out;
    // 6-point Gauss-Legendre on [sa, sb]
    const double gx[kNs] = {-0.932469514203152, -0.661209386466265, -0.238619186083197,
                            0.238619186083197, 0.661209386466265, 0.932469514203152};
    const double gw[kNs] = {0.171324492379170, 0.360761573048139, 0.467913934572691,
                            0.467913934572691, 0.360761573048139, 0.171324492379170};
    const double s = 0.5 * (sa + sb) + 0.5 * (sb - sa) * gx[j];
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
    out.weight = (float)(0.5 * (sb - sa) * gw[j] * norm);
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
    out.piece = piece;
    return out;
}

__device__ __forceinline__ Slot shuffle(const Slot& s, int lane) {
    Slot o;
    o.weight = __shfl(s.weight, lane);
    o.D = __shfl(s.D, lane);
    o.D2R2 = __shfl(s.D2R2, lane);
    o.eps = __shfl(s.eps, lane);
    o.hav_eps = __shfl(s.hav_eps, lane);
    o.sinpsi = __shfl(s.sinpsi, lane);
    o.psi_theta_a = __shfl(s.psi_theta_a, lane);
    o.theta_a = __shfl(s.theta_a, lane);
    o.cos_theta_a = __shfl(s.cos_theta_a, lane);
    o.ta = __shfl(s.ta, lane);
    o.tb = __shfl(s.tb, lane);
    o.piece = __shfl(s.piece, lane);
    return o;
}

struct PanelTables {
    // medium panels between consecutive major lines
    float lam_a[kMaxMajor], dlam_dn[kMaxMajor], n_a[kMaxMajor];
    float mu_a[kMaxMajor], dmu[kMaxMajor], span[kMaxMajor];
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

// Value of the lane's (theta, phi) node of one slot (without the slot weight).
__device__ float node_value(const PanelTables& T, const Slot& S, int npanel, int nalpha,
                            float beta, float c0, float c2, float inv_length, float axial,
                            float across, float side, float xt, float wt, float se, float ce,
                            float we) {
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
    const float dn = 2.f * sinf(S.theta_a + 0.5f * dth) * sinf(0.5f * dth) / (beta * cth * S.cos_theta_a);
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
    const float H = sinf(0.5f * S.eps * om) * sinf(0.5f * S.eps * op);
    const float sts = sth * S.sinpsi;
    const float hh = fminf(H / sts, 1.f);
    if (!(hh > 0.f)) return 0.f;
    const float half = 2.f * asinf(sqrtf(hh));
    const float phi = half * se;
    const float dphi = half * ce * we;
    const float hphi = sinf(0.5f * phi);
    const float hdel = sinf(0.5f * delta);
    const float hav_g = hdel * hdel + sts * hphi * hphi;
    // hav(eps) - hav(gamma) = sin(theta) sin(psi) sin((half - phi)/2) sin((half + phi)/2)
    float em, ep;
    one_minus_plus(se, ce, em, ep);
    const float gap = sts * sinf(0.5f * half * em) * sinf(0.5f * half * ep);
    const float inside = fmaxf(gap * (1.f - S.hav_eps - hav_g), 0.f);
    const float d = S.D2R2 / (S.D * (1.f - 2.f * hav_g) + 2.f * S.D * sqrtf(inside));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float acc = acceptance(T.alpha, nalpha, cth * axial + sth * (cp * across + sp * side))
                    + acceptance(T.alpha, nalpha, cth * axial + sth * (cp * across - sp * side));
    return hw * wt * S.eps * ct * spectral * jac * dphi * acc * expf(-mu * d);
}

__global__ __launch_bounds__(kThreads) void order0_charge(O0Params P) {
    __shared__ PanelTables T;
    __shared__ int list[2048];
    __shared__ int wave_count[kWaves];
    __shared__ int total_count;
    __shared__ double wave_sum[kWaves];
    __shared__ int glass;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int chunk = blockIdx.x;
    const int om = blockIdx.y;

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
    }
    for (int k = tid; k <= P.npanel; k += kThreads) T.edges[k] = (float)P.edges[k];
    for (int k = tid; k < P.npanel * 4; k += kThreads)
        T.eff[k / 4][k % 4] = (float)P.eff[(size_t)(k / 4) * P.effcols + (k % 4)];
    if (tid < P.nalpha) T.alpha[tid] = (float)(P.alpha[tid] / (4.0 * kPi));
    if (tid == 0) {
        total_count = 0;
        glass = 0;
    }
    __syncthreads();

    // candidate pairs of this chunk, compacted in element order
    const int first = chunk * P.chunk;
    const int last = min(P.nel, first + P.chunk);
    int local_glass = 0;
    for (int base = first; base < last; base += kThreads) {
        const int i = base + tid;
        const int pass = (i < last) ? candidate(P, om, i, &local_glass) : 0;
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

    // integrate: one wave per pair
    const int it = lane >> 3;            // theta node
    const int ie = lane & 7;             // azimuth node
    const float gxt[kNt] = {-0.960289856497536f, -0.796666477413627f, -0.525532409916329f,
                            -0.183434642495650f, 0.183434642495650f, 0.525532409916329f,
                            0.796666477413627f, 0.960289856497536f};
    const float gwt[kNt] = {0.101228536290376f, 0.222381034453374f, 0.313706645877887f,
                            0.362683783378362f, 0.362683783378362f, 0.313706645877887f,
                            0.222381034453374f, 0.101228536290376f};
    const float xt = gxt[it], wt = gwt[it];
    // azimuth node on [0, pi/2]: eta = pi/4 (x + 1)
    const float eta_node = 0.78539816339744831f * (gxt[ie] + 1.f);
    const float we = 0.78539816339744831f * gwt[ie];
    float se, ce;
    sincosf(eta_node, &se, &ce);

    const double* pos = P.positions + 3 * om;
    const double* look = P.looks + 3 * om;
    const double R = P.radii[om];
    const double norm = P.areas[om] / (kPi * R * R) / (2.0 * kPi);
    double lane_sum = 0.0;
    for (int entry = wave; entry < count; entry += kWaves) {
        const int i = list[entry];
        const double ux = P.dir[3 * i], uy = P.dir[3 * i + 1], uz = P.dir[3 * i + 2];
        const double wx = pos[0] - P.start[3 * i], wy = pos[1] - P.start[3 * i + 1],
                     wz = pos[2] - P.start[3 * i + 2];
        const double z = wx * ux + wy * uy + wz * uz;
        const double tx = wx - z * ux, ty = wy - z * uy, tz = wz - z * uz;
        const double rho = sqrt(tx * tx + ty * ty + tz * tz);
        if (!(rho > R)) {
            if (lane == 0) atomicOr(&P.flags[0], 2);
            continue;
        }
        const double e1x = tx / rho, e1y = ty / rho, e1z = tz / rho;
        const double e2x = uy * e1z - uz * e1y, e2y = uz * e1x - ux * e1z, e2z = ux * e1y - uy * e1x;
        const float axial = (float)(ux * look[0] + uy * look[1] + uz * look[2]);
        const float across = (float)(e1x * look[0] + e1y * look[1] + e1z * look[2]);
        const float side = (float)(e2x * look[0] + e2y * look[1] + e2z * look[2]);
        const double L = P.length[i], beta = P.beta[i];
        // cone angles of the medium panel edges and the s pieces
        double theta_major[kMaxMajor];
        double scut[2 * kMaxMajor + 2];
        int nscut = 0;
        scut[nscut++] = 0.0;
        for (int m = 0; m < P.nmajor; ++m) {
            const double c = 1.0 / (beta * P.phase[P.major[m]]);
            const double sn = sqrt(fmax(0.0, (1.0 - c) * (1.0 + c)));
            theta_major[m] = atan2(sn, c);
            if (sn > 0.0) {
                const double centre = z - rho * c / sn, spread = R / sn;
                const double r1 = centre - spread, r2 = centre + spread;
                if (r1 > 0.0 && r1 < L) scut[nscut++] = r1;
                if (r2 > 0.0 && r2 < L) scut[nscut++] = r2;
            }
        }
        scut[nscut++] = L;
        for (int a = 1; a < nscut; ++a)  // insertion sort (few entries)
            for (int b = a; b > 0 && scut[b - 1] > scut[b]; --b) {
                const double tmp = scut[b];
                scut[b] = scut[b - 1];
                scut[b - 1] = tmp;
            }
        const int slots = (nscut - 1) * kNs * pieces;
        const float c0 = (float)P.c0[i], c2 = (float)P.c2[i];
        const float inv_length = (float)(1.0 / L);
        float pair_sum = 0.f;
        for (int group = 0; group < slots; group += 64) {
            const Slot mine = (group + lane < slots)
                ? make_slot(P, group + lane, pieces, scut, nscut, z, rho, L, R, beta, theta_major, norm)
                : Slot{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1};
            const int active = min(64, slots - group);
            for (int k = 0; k < active; ++k) {
                const Slot S = shuffle(mine, k);
                if (S.piece < 0) continue;
                pair_sum += S.weight * node_value(T, S, P.npanel, P.nalpha, (float)beta, c0, c2,
                                                  inv_length, axial, across, side, xt, wt, se, ce, we);
            }
        }
        lane_sum += (double)pair_sum;
    }
    // fixed-order reductions: lanes, then waves
    for (int offset = 32; offset > 0; offset >>= 1) lane_sum += __shfl_down(lane_sum, offset);
    if (lane == 0) wave_sum[wave] = lane_sum;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
        for (int w = 0; w < kWaves; ++w) sum += wave_sum[w];
        P.partial[(size_t)om * P.nchunk + chunk] = sum;
    }
}

__global__ void order0_reduce(const double* partial, int nom, int nchunk, double* charge) {
    const int om = blockIdx.x * blockDim.x + threadIdx.x;
    if (om >= nom) return;
    double sum = 0.0;
    for (int c = 0; c < nchunk; ++c) sum += partial[(size_t)om * nchunk + c];
    charge[om] = sum;
}

constexpr int kChunk = 2048;

}  // namespace

extern "C" size_t lh_order0_workspace_bytes(int n_oms, int n_elements) {
    const size_t nchunk = n_elements > 0 ? (size_t)(n_elements + kChunk - 1) / kChunk : 1;
    return 256 + sizeof(double) * (size_t)(n_oms > 0 ? n_oms : 0) * nchunk;
}

extern "C" int lh_order0_charge(const lh_order0_optics* optics, const lh_order0_oms* oms,
                                const lh_order0_elements* elements, double* charge,
                                void* workspace, size_t workspace_bytes, void* stream) {
    LH_REQUIRE(optics && oms && elements, "lh_order0_charge: null descriptor");
    LH_REQUIRE(oms->count >= 0 && elements->count >= 0, "lh_order0_charge: negative count");
    LH_REQUIRE(workspace_bytes >= lh_order0_workspace_bytes(oms->count, elements->count),
               "lh_order0_charge: workspace of %zu bytes, %zu needed", workspace_bytes,
               lh_order0_workspace_bytes(oms->count, elements->count));
    LH_REQUIRE(optics->panels >= 1 && optics->panels <= kMaxPanels,
               "lh_order0_charge: 1..%d optical panels supported", kMaxPanels);
    LH_REQUIRE(optics->major_count >= 2 && optics->major_count <= kMaxMajor,
               "lh_order0_charge: 2..%d medium nodes in the band supported", kMaxMajor);
    LH_REQUIRE(optics->acceptance_count >= 1 && optics->acceptance_count <= kMaxAlpha,
               "lh_order0_charge: 1..%d acceptance coefficients supported", kMaxAlpha);
    LH_REQUIRE(optics->efficiency_columns == 4,
               "lh_order0_charge: cubic efficiency panels (4 columns) required");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* flags = static_cast<int*>(workspace);
    double* partial = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256);
    LH_HIP_CALL(hipMemsetAsync(flags, 0, 256, s));
    if (oms->count == 0) return LH_OK;
    if (elements->count == 0) {
        LH_HIP_CALL(hipMemsetAsync(charge, 0, sizeof(double) * oms->count, s));
        return LH_OK;
    }
    O0Params P{};
    P.positions = oms->positions;
    P.looks = oms->looks;
    P.areas = oms->areas;
    P.radii = oms->radii;
    P.nom = oms->count;
    P.start = elements->start_m;
    P.dir = elements->direction;
    P.length = elements->length_m;
    P.c0 = elements->coefficient0;
    P.c2 = elements->coefficient2;
    P.beta = elements->beta;
    P.nel = elements->count;
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
    P.chunk = kChunk;
    P.nchunk = (elements->count + kChunk - 1) / kChunk;
    P.partial = partial;
    P.flags = flags;
    hipLaunchKernelGGL(order0_charge, dim3(P.nchunk, P.nom), dim3(kThreads), 0, s, P);
    int status = lh::check_launch("lh_order0_charge");
    if (status != LH_OK) return status;
    hipLaunchKernelGGL(order0_reduce, dim3((P.nom + 255) / 256), dim3(256), 0, s, partial, P.nom,
                       P.nchunk, charge);
    return lh::check_launch("lh_order0_charge (reduce)");
}
