// Order >= 2 source: axial-slice compile of one light field (gfx950).
//
// Restates compile_axial_slices (_shower_basis.py:188-645) for the shower
// route's options. Lattice: every chord node links to (slice, tick) rows with
// weight wz * wt; rows are the occupied keys slice * kmax + tick of a dense
// occupancy map. The reference deposits rows with dense products over groups
// of nodes sharing a base key; here the nodes are radix-sorted by base key and
// every row gathers its (at most eight) base-key buckets in sorted order, so
// each sum has one fixed order (deterministic, no atomics). The slice
// geometry (thin/wide, centroid) is summed from the same buckets with the
// axial weights alone, as in the reference (:373-396). Per slice, the rows are
// summed with e^{i omega tau dt} in FP64 and combined into the five-point
// stencil (:563-598). All arithmetic is FP64; channels are written complex64.
#include "lh_common.h"
#include "lh_sort.h"

#include <hipcub/hipcub.hpp>

#include <cmath>

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kMaxDegree = 16;
constexpr int kMaxFreq = 128;
constexpr int kLayoutThreads = 1024;

// info words written by lh_compile_layout
enum { kInfoSlices = 0, kInfoWide, kInfoPoints, kInfoRows, kInfoFlags, kInfoWords = 8 };
constexpr int kFlagRange = 1;      // a node fell outside the lattice

struct NodeRecord {
    double wz[4];   // axial weights of the node's slice shifts
    double tr;      // fractional tick (linear lattice)
    double f[5];    // photons * moment * node weight
};

struct Params {
    lh_compile_field field;
    lh_compile_lattice lat;
    int nodes;        // chord nodes per element
    int channels;
    int shifts;       // axial links per node (2 or 4)
    int shift_lo;     // first shift (-1 cubic, 0 linear)
    int ticks;        // tick links per node (1 or 2)
    long long span;   // positions * kmax
    double node_frac[2], node_weight[2];
    // workspace views
    double* angular;          // (count, channels)
    NodeRecord* records;      // (nodes * count)
    int* keys;                // (nodes * count) base keys, span for skipped nodes
    int* ids;
    int* sorted_keys;
    int* sorted_ids;
    int* occupied;            // (span + 1)
    int* row_scan;            // (span + 1) exclusive
    int* row_key;             // (span)
    int* slice_used;          // (positions)
    double* zero;             // (positions, 5)
    int* wide;                // (positions)
    int* point_of;            // (positions) index of the slice's first point
    int* wide_of;             // (positions) index among the wide slices
    int* info;
    double* points;           // (5 * positions, 3)
};

__global__ void __launch_bounds__(kThreads) compile_clear(int* a, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__device__ inline bool element_taken(const lh_compile_field& f, int i) {
    if (f.mask && !f.mask[i]) return false;
    if (f.photons) return true;
    return f.beta[i] * f.phase_index > 1.0;
}

__device__ inline double element_photons(const lh_compile_field& f, int i) {
    if (f.photons) return f.photons[i];
    double first = f.coefficient0[i];
    double second = -f.coefficient2[i];
    double a = f.scale0 * first;
    double b = f.scale2 * second;
    return a - b;
}

__device__ inline double element_cone(const lh_compile_field& f, int i) {
    if (f.cone_cosine) return f.cone_cosine[i];
    return 1.0 / (f.beta[i] * f.phase_index);
}

// P_l(cone) * Y_lm(u) of every kept channel (axial_fast.py:21-52,
// _shower_basis.py:82-100), u = direction in the frame.
__global__ void __launch_bounds__(kThreads) compile_angular(Params p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.field.count) return;
    const lh_compile_field& f = p.field;
    const lh_compile_lattice& L = p.lat;
    double* out = p.angular + (size_t)i * p.channels;
    if (!element_taken(f, i)) {
        for (int c = 0; c < p.channels; ++c) out[c] = 0.0;
        return;
    }
    double d0 = f.direction[3 * i], d1 = f.direction[3 * i + 1], d2 = f.direction[3 * i + 2];
    double x = d0 * L.first[0] + d1 * L.first[1] + d2 * L.first[2];
    double y = d0 * L.second[0] + d1 * L.second[1] + d2 * L.second[2];
    double z = d0 * L.axis[0] + d1 * L.axis[1] + d2 * L.axis[2];
    int degree = L.degree;
    int max_m = min(L.degree, L.azimuthal_degree);
    double cone = element_cone(f, i);
    double legendre[kMaxDegree + 1];
    legendre[0] = 1.0;
    if (degree >= 1) legendre[1] = cone;
    for (int l = 2; l <= degree; ++l)
        legendre[l] = ((2 * l - 1) * cone * legendre[l - 1] - (l - 1) * legendre[l - 2]) / l;
    double radius = sqrt(x * x + y * y + z * z);
    double mu = z / radius;
    double transverse = sqrt(x * x + y * y);
    double sine = transverse / radius;
    double cp = transverse == 0.0 ? 1.0 : x / transverse;
    double sp = transverse == 0.0 ? 0.0 : y / transverse;
    double cm = 1.0, sm = 0.0;
    double diagonal = 1.0 / sqrt(4 * M_PI);
    const double root2 = sqrt(2.0);
    for (int m = 0; m <= max_m; ++m) {
        if (m) {
            diagonal *= sqrt((2.0 * m + 1) / (2.0 * m)) * sine;
            double c2 = cm * cp - sm * sp;
            double s2 = sm * cp + cm * sp;
            cm = c2;
            sm = s2;
        }
        double v0 = diagonal, v1 = 0.0;
        for (int l = m; l <= degree; ++l) {
            double value;
            if (l == m) {
                value = v0;
            } else if (l == m + 1) {
                value = sqrt(2.0 * m + 3) * mu * v0;
                v1 = value;
            } else {
                double a = sqrt((4.0 * l * l - 1) / (double)(l * l - m * m));
                double b = sqrt(((2.0 * l + 1) * ((l - 1) * (l - 1) - m * m))
                                / ((2.0 * l - 3) * (double)(l * l - m * m)));
                value = a * mu * v1 - b * v0;
                v0 = v1;
                v1 = value;
            }
            // kept channels of degree l start at sum_{k<l} (2 min(k, M) + 1)
            int start = 0;
            for (int k = 0; k < l; ++k) start += 2 * min(k, max_m) + 1;
            int centre = start + min(l, max_m);
            if (m == 0) {
                out[centre] = value * legendre[l];
            } else {
                out[centre - m] = root2 * value * sm * legendre[l];
                out[centre + m] = root2 * value * cm * legendre[l];
            }
        }
    }
}

__device__ inline void axial_weights(double r, int order, double w[4]) {
    if (order == 1) {
        w[0] = 1 - r;
        w[1] = r;
        w[2] = w[3] = 0.0;
        return;
    }
    w[0] = -r * (r - 1) * (r - 2) / 6;
    w[1] = (r + 1) * (r - 1) * (r - 2) / 2;
    w[2] = -(r + 1) * r * (r - 2) / 2;
    w[3] = (r + 1) * r * (r - 1) / 6;
}

// Chord node -> base key, weights, moments; marks the rows and slices it links to
// (_shower_basis.py:316-355).
__global__ void __launch_bounds__(kThreads) compile_nodes(Params p) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    int count = p.field.count;
    if (g >= p.nodes * count) return;
    int n = g / count, i = g - n * count;
    const lh_compile_field& f = p.field;
    const lh_compile_lattice& L = p.lat;
    p.ids[g] = g;
    if (!element_taken(f, i)) {
        p.keys[g] = (int)p.span;
        return;
    }
    double frac = p.node_frac[n];
    double len = frac * f.length_m[i];
    double pt[3], offset[3];
    for (int k = 0; k < 3; ++k) {
        pt[k] = f.start_m[3 * i + k] + len * f.direction[3 * i + k];
        offset[k] = pt[k] - L.centre[k];
    }
    double z = offset[0] * L.axis[0] + offset[1] * L.axis[1] + offset[2] * L.axis[2];
    double tv[3];
    for (int k = 0; k < 3; ++k) tv[k] = offset[k] - z * L.axis[k];
    double x = tv[0] * L.first[0] + tv[1] * L.first[1] + tv[2] * L.first[2];
    double y = tv[0] * L.second[0] + tv[1] * L.second[1] + tv[2] * L.second[2];
    double t0 = f.start_ns[i];
    double time = t0 + frac * (f.end_ns[i] - t0);
    NodeRecord rec;
    double photons = element_photons(f, i);
    double w = p.node_weight[n];
    double mom[5] = {1.0, x, y, x * x, y * y};
    for (int k = 0; k < 5; ++k) rec.f[k] = k < L.moments ? (photons * mom[k]) * w : 0.0;
    double zl = (z - L.z_origin_m) / L.cell_m;
    double zbf = floor(zl);
    axial_weights(zl - zbf, L.axial_order, rec.wz);
    long long zb = (long long)zbf;
    long long tb = 0;
    rec.tr = 0.0;
    if (L.time_deposit != 2) {
        double tl = (time - L.reference_ns) / L.time_step_ns;
        if (L.time_deposit == 1) {
            double tbf = floor(tl);
            tb = (long long)tbf;
            rec.tr = tl - tbf;
        } else {
            tb = (long long)rint(tl);
        }
    }
    long long base = zb - L.slice0;
    bool inside = tb >= 0 && tb + p.ticks - 1 < L.kmax && base + p.shift_lo >= 0
                  && base + p.shift_lo + p.shifts - 1 < L.positions;
    if (!inside) {
        atomicOr(p.info + kInfoFlags, kFlagRange);
        p.keys[g] = (int)p.span;
        return;
    }
    p.records[g] = rec;
    p.keys[g] = (int)(base * L.kmax + tb);
    for (int j = 0; j < p.shifts; ++j) {
        long long s = base + p.shift_lo + j;
        p.slice_used[s] = 1;
        for (int dt = 0; dt < p.ticks; ++dt) p.occupied[s * L.kmax + tb + dt] = 1;
    }
}

__device__ inline int lower_bound(const int* keys, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Zero-frequency moments of every slice from the axial weights alone and the
// thin/wide decision (_shower_basis.py:373-396, stencil_center "axis").
__global__ void __launch_bounds__(kThreads) compile_geometry(Params p) {
    __shared__ double part[5][kThreads];
    __shared__ int range[4][2];
    int s = blockIdx.x;
    const lh_compile_lattice& L = p.lat;
    int total = p.nodes * p.field.count;
    if (!p.slice_used[s]) {
        if (threadIdx.x == 0) {
            for (int k = 0; k < 5; ++k) p.zero[5 * s + k] = 0.0;
            p.wide[s] = 0;
        }
        return;
    }
    if (threadIdx.x < p.shifts) {
        int j = threadIdx.x;
        long long b = (long long)s - (p.shift_lo + j);
        int lo = 0, hi = 0;
        if (b >= 0 && b < L.positions) {
            lo = lower_bound(p.sorted_keys, total, (int)(b * L.kmax));
            hi = lower_bound(p.sorted_keys, total, (int)((b + 1) * L.kmax));
        }
        range[j][0] = lo;
        range[j][1] = hi;
    }
    __syncthreads();
    double acc[5] = {0, 0, 0, 0, 0};
    for (int j = 0; j < p.shifts; ++j) {
        for (int q = range[j][0] + threadIdx.x; q < range[j][1]; q += kThreads) {
            const NodeRecord& r = p.records[p.sorted_ids[q]];
            double w = r.wz[j];
            for (int k = 0; k < L.moments; ++k) acc[k] += w * r.f[k];
        }
    }
    for (int k = 0; k < 5; ++k) part[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half)
            for (int k = 0; k < 5; ++k) part[k][threadIdx.x] += part[k][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) p.zero[5 * s + k] = part[k][0];
        int wide = 0;
        if (L.moments == 5) {
            double mass = part[0][0];
            double safe = fabs(mass) > 0 ? mass : 1.0;
            double spread = sqrt(fabs(part[3][0] + part[4][0]) / fabs(safe));
            wide = spread > L.stencil_threshold_m;
        }
        p.wide[s] = wide;
    }
}

// Block-wide exclusive scan of 0/1 flags over positions (one workgroup).
__device__ int block_scan(const int* flags, int* out, int n, int* smem) {
    int carry = 0;
    for (int base = 0; base < n; base += kLayoutThreads) {
        int i = base + threadIdx.x;
        int v = i < n ? (flags[i] != 0) : 0;
        smem[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kLayoutThreads; off <<= 1) {
            int add = threadIdx.x >= off ? smem[threadIdx.x - off] : 0;
            __syncthreads();
            smem[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) out[i] = carry + smem[threadIdx.x] - v;
        int block_total = smem[kLayoutThreads - 1];
        __syncthreads();
        carry += block_total;
    }
    return carry;
}

// Point layout: one point per slice, then +x, -x, +y, -y of the wide slices
// (_shower_basis.py:446-455, 599-614). Writes the info words.
__global__ void __launch_bounds__(kLayoutThreads) compile_layout(Params p) {
    __shared__ int smem[kLayoutThreads];
    const lh_compile_lattice& L = p.lat;
    int slices = block_scan(p.slice_used, p.point_of, L.positions, smem);
    int wide = block_scan(p.wide, p.wide_of, L.positions, smem);
    double h = L.stencil_m;
    for (int s = threadIdx.x; s < L.positions; s += kLayoutThreads) {
        if (!p.slice_used[s]) continue;
        double zc = L.z_origin_m + (double)(s + L.slice0) * L.cell_m;
        int main = p.point_of[s];
        double cx = 0.0, cy = 0.0;
        if (L.moments == 5 && !p.wide[s]) {
            double mass = p.zero[5 * s];
            double safe = fabs(mass) > 0 ? mass : 1.0;
            cx = p.zero[5 * s + 1] / safe;
            cy = p.zero[5 * s + 2] / safe;
        }
        double* q = p.points + 3 * (size_t)main;
        q[0] = cx; q[1] = cy; q[2] = zc;
        if (p.wide[s]) {
            const double ox[4] = {h, -h, 0.0, 0.0};
            const double oy[4] = {0.0, 0.0, h, -h};
            for (int o = 0; o < 4; ++o) {
                double* e = p.points + 3 * (size_t)(slices + o * wide + p.wide_of[s]);
                e[0] = ox[o]; e[1] = oy[o]; e[2] = zc;
            }
        }
    }
    if (threadIdx.x == 0) {
        p.info[kInfoSlices] = slices;
        p.info[kInfoWide] = wide;
        p.info[kInfoPoints] = slices + 4 * wide;
        p.info[kInfoRows] = p.row_scan[p.span];
    }
}

__global__ void __launch_bounds__(kThreads) compile_row_keys(Params p) {
    long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= p.span) return;
    if (p.occupied[key]) p.row_key[p.row_scan[key]] = (int)key;
}

// Deposit of one lattice row: sum over its base-key buckets in sorted order of
// (wz * wt * moment) * P_l Y_lm (_shower_basis.py:482-555).
//
// The bucket's nodes are staged through LDS in tiles of `tile` nodes: the
// products (wz * wt) * moment of every node (5 per node) and its angular row
// (channels values, coalesced), so each node's record and angular row are
// read from memory once per row instead of once per (moment, channel) thread.
// Every thread then adds its (moment, channel) values in node order with the
// same rounding as a direct loop.
constexpr int kDepositAcc = (5 * (kMaxDegree + 1) * (kMaxDegree + 1) + kThreads - 1) / kThreads;

// Nodes per LDS tile of compile_deposit for `channels` angular channels.
inline int deposit_tile(int channels) {
    return std::max(8, std::min(64, 32768 / ((channels + 6) * 8)));
}

inline size_t deposit_lds_bytes(int channels) {
    const int tile = deposit_tile(channels);
    return (size_t)tile * (5 + channels) * sizeof(double) + (size_t)tile * sizeof(int);
}

__global__ void __launch_bounds__(kThreads) compile_deposit(Params p, int rows, int tile,
                                                            double* deposit) {
    extern __shared__ double stage[];   // (tile, 5) products, (tile, channels), (tile) ids
    __shared__ int range[8][2];
    __shared__ int link[8][2];   // shift index, tick offset
    int r = blockIdx.x;
    if (r >= rows) return;
    const lh_compile_lattice& L = p.lat;
    int total = p.nodes * p.field.count;
    int key = p.row_key[r];
    int s = key / L.kmax, tau = key - s * L.kmax;
    int links = p.shifts * p.ticks;
    if (threadIdx.x < links) {
        int j = threadIdx.x / p.ticks, dt = threadIdx.x - j * p.ticks;
        long long b = (long long)s - (p.shift_lo + j);
        int t0 = tau - dt;
        int lo = 0, hi = 0;
        if (b >= 0 && b < L.positions && t0 >= 0) {
            int bk = (int)(b * L.kmax + t0);
            lo = lower_bound(p.sorted_keys, total, bk);
            hi = lower_bound(p.sorted_keys, total, bk + 1);
        }
        range[threadIdx.x][0] = lo;
        range[threadIdx.x][1] = hi;
        link[threadIdx.x][0] = j;
        link[threadIdx.x][1] = dt;
    }
    __syncthreads();
    const int C = p.channels;
    const int width = L.moments * C;
    const int count = p.field.count;
    double* s_wf = stage;
    double* s_ang = stage + 5 * tile;
    int* s_id = reinterpret_cast<int*>(stage + (size_t)(5 + C) * tile);
    double acc[kDepositAcc];
#pragma unroll
    for (int a = 0; a < kDepositAcc; ++a) acc[a] = 0.0;
    for (int l = 0; l < links; ++l) {
        const int j = link[l][0], dt = link[l][1];
        const int q_hi = range[l][1];
        for (int q0 = range[l][0]; q0 < q_hi; q0 += tile) {
            const int n = min(tile, q_hi - q0);
            __syncthreads();   // the previous tile has been read
            if ((int)threadIdx.x < n) {
                const int id = p.sorted_ids[q0 + threadIdx.x];
                const NodeRecord& rec = p.records[id];
                double wt = L.time_deposit == 1 ? (dt ? rec.tr : 1 - rec.tr) : 1.0;
                double w = rec.wz[j] * wt;
#pragma unroll
                for (int k = 0; k < 5; ++k) s_wf[threadIdx.x * 5 + k] = w * rec.f[k];
                s_id[threadIdx.x] = id % count;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < n * C; e += kThreads) {
                const int i = e / C, c = e - i * C;
                s_ang[e] = p.angular[(size_t)s_id[i] * C + c];
            }
            __syncthreads();
#pragma unroll
            for (int a = 0; a < kDepositAcc; ++a) {
                const int kc = threadIdx.x + a * kThreads;
                if (kc < width) {
                    const int k = kc / C, c = kc - k * C;
                    double sum = acc[a];
                    for (int i = 0; i < n; ++i) sum += s_wf[i * 5 + k] * s_ang[i * C + c];
                    acc[a] = sum;
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < kDepositAcc; ++a) {
        const int kc = threadIdx.x + a * kThreads;
        if (kc < width) deposit[(size_t)r * width + kc] = acc[a];
    }
}

// Per slice: values[k, c, w] = sum_rows D[row, k, c] e^{i omega_w tau dt}, then the
// thin point or the five stencil points (_shower_basis.py:563-598).
// Phase table of every row: e^{i omega t} at the row's tick, t = tick * dt (FP64).
__global__ void __launch_bounds__(kThreads) compile_phase_table(Params p, const double* omega,
                                                                int nfreq, int rows,
                                                                double* table_c,
                                                                double* table_s) {
    long long e = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (long long)rows * nfreq) return;
    int r = (int)(e / nfreq), w = (int)(e - (long long)r * nfreq);
    const lh_compile_lattice& L = p.lat;
    int key = p.row_key[r];
    double t = (double)(key % L.kmax) * L.time_step_ns;
    double sv, cv;
    sincos(t * omega[w], &sv, &cv);
    table_c[e] = cv;
    table_s[e] = sv;
}

// Per slice and channel: the rows' deposits summed with the phase table in row
// order (FP64), then the stencil combination. One wave per (slice, channel);
// lane w holds frequencies w and w + 64; the deposits of a row are the same
// for every lane.
__global__ void __launch_bounds__(64) compile_phase(Params p, int nfreq, double scale,
                                                   const double* deposit, const double* table_c,
                                                   const double* table_s, float2* channels) {
    int s = blockIdx.x;
    if (!p.slice_used[s]) return;
    const lh_compile_lattice& L = p.lat;
    const int c = blockIdx.y;
    const int lane = threadIdx.x;
    const int width = L.moments * p.channels;
    const int row_lo = p.row_scan[(long long)s * L.kmax];
    const int row_hi = p.row_scan[(long long)(s + 1) * L.kmax];
    double re[2][5], im[2][5];
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 5; ++k) re[b][k] = im[b][k] = 0.0;
    const bool second = lane + 64 < nfreq;
    const bool first = lane < nfreq;
    // rows four at a time: every load of the four is issued before the sums
    // that use them, in row order as before
    constexpr int kRows = 4;
    const int moments = L.moments;
    for (int r0 = row_lo; r0 < row_hi; r0 += kRows) {
        double c0[kRows], s0[kRows], c1[kRows], s1[kRows], v[kRows][5];
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            const int r = r0 + i;
            const bool in = r < row_hi;
            const double* d = deposit + (size_t)r * width + c;
            const double* tc = table_c + (size_t)r * nfreq;
            const double* ts = table_s + (size_t)r * nfreq;
            c0[i] = in && first ? tc[lane] : 0.0;
            s0[i] = in && first ? ts[lane] : 0.0;
            c1[i] = in && second ? tc[lane + 64] : 0.0;
            s1[i] = in && second ? ts[lane + 64] : 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) v[i][k] = in && k < moments ? d[k * p.channels] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            if (r0 + i >= row_hi) break;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                if (k >= moments) break;
                re[0][k] += v[i][k] * c0[i];
                im[0][k] += v[i][k] * s0[i];
                re[1][k] += v[i][k] * c1[i];
                im[1][k] += v[i][k] * s1[i];
            }
        }
    }
    int slices = p.info[kInfoSlices], wide_count = p.info[kInfoWide];
    double h = L.stencil_m;
    for (int b = 0; b < 2; ++b) {
        int w = lane + 64 * b;
        if (w >= nfreq) continue;
        if (L.moments == 1 || !p.wide[s]) {
            size_t at = ((size_t)p.point_of[s] * p.channels + c) * nfreq + w;
            channels[at] = make_float2((float)(re[b][0] * scale), (float)(im[b][0] * scale));
            continue;
        }
        double h2 = h * h;
        double vr[5], vi[5];
        double m0r = re[b][0], mxr = re[b][1], myr = re[b][2];
        double mxxr = re[b][3], myyr = re[b][4];
        double m0i = im[b][0], mxi = im[b][1], myi = im[b][2];
        double mxxi = im[b][3], myyi = im[b][4];
        vr[0] = m0r - (mxxr + myyr) / h2;         vi[0] = m0i - (mxxi + myyi) / h2;
        vr[1] = 0.5 * (mxxr / h2 + mxr / h);      vi[1] = 0.5 * (mxxi / h2 + mxi / h);
        vr[2] = 0.5 * (mxxr / h2 - mxr / h);      vi[2] = 0.5 * (mxxi / h2 - mxi / h);
        vr[3] = 0.5 * (myyr / h2 + myr / h);      vi[3] = 0.5 * (myyi / h2 + myi / h);
        vr[4] = 0.5 * (myyr / h2 - myr / h);      vi[4] = 0.5 * (myyi / h2 - myi / h);
        int ids[5] = {p.point_of[s], slices + p.wide_of[s], slices + wide_count + p.wide_of[s],
                      slices + 2 * wide_count + p.wide_of[s],
                      slices + 3 * wide_count + p.wide_of[s]};
        for (int o = 0; o < 5; ++o) {
            size_t at = ((size_t)ids[o] * p.channels + c) * nfreq + w;
            channels[at] = make_float2((float)(vr[o] * scale), (float)(vi[o] * scale));
        }
    }
}

int bits_for(long long value) {
    int bits = 1;
    while (bits < 31 && (1LL << bits) <= value) ++bits;
    return bits;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
    size_t angular, records, keys, ids, sorted_keys, sorted_ids, occupied, row_scan, row_key;
    size_t slice_used, zero, wide, point_of, wide_of, info, sort_temp, scan_temp, total;
};

int channel_count(int degree, int max_m) {
    int n = 0;
    for (int l = 0; l <= degree; ++l) n += 2 * std::min(l, max_m) + 1;
    return n;
}

int check_lattice(const lh_compile_lattice* L, int count) {
    LH_REQUIRE(L != nullptr, "lattice descriptor is NULL");
    LH_REQUIRE(count > 0, "the field needs at least one element");
    LH_REQUIRE(L->degree >= 0 && L->degree <= kMaxDegree, "source degree must lie in 0..%d",
               kMaxDegree);
    LH_REQUIRE(L->azimuthal_degree >= 0, "azimuthal degree must be nonnegative");
    LH_REQUIRE(L->axial_order == 1 || L->axial_order == 3, "axial_order must be 1 or 3");
    LH_REQUIRE(L->moments == 1 || L->moments == 5, "moments must be 1 (collapse) or 5 (stencil)");
    LH_REQUIRE(L->element_order == 1 || L->element_order == 2, "element_order must be 1 or 2");
    LH_REQUIRE(L->time_deposit >= 0 && L->time_deposit <= 2, "time_deposit must be 0, 1 or 2");
    LH_REQUIRE(L->kmax >= 1 && L->positions >= 1, "empty lattice");
    LH_REQUIRE((long long)L->kmax * L->positions < (1LL << 30),
               "source too extended for the dense slice lattice (%d slices x %d ticks)",
               L->positions, L->kmax);
    LH_REQUIRE(L->cell_m > 0 && L->time_step_ns > 0 && L->stencil_m > 0
                   && L->stencil_threshold_m >= 0,
               "cell_m, time_step_ns and stencil_m must be positive");
    LH_REQUIRE((long long)count * L->element_order < (1LL << 30), "too many chord nodes");
    return LH_OK;
}

int make_layout(const lh_compile_lattice* L, int count, Layout& out) {
    int nodes = L->element_order;
    size_t total_nodes = (size_t)nodes * count;
    size_t span = (size_t)L->kmax * L->positions;
    int channels = channel_count(L->degree, std::min(L->degree, L->azimuthal_degree));
    size_t sort_temp = 0, scan_temp = 0;
    if (lh::sort_pairs(nullptr, sort_temp, (unsigned*)nullptr,
                                           (unsigned*)nullptr, (int*)nullptr, (int*)nullptr,
                                           (int)total_nodes, 0, bits_for((long long)span),
                                           (hipStream_t)0) != hipSuccess) {
        lh::set_error("radix sort workspace query failed");
        return LH_ERR_HIP;
    }
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_temp, (int*)nullptr, (int*)nullptr,
                                         (int)(span + 1), (hipStream_t)0) != hipSuccess) {
        lh::set_error("scan workspace query failed");
        return LH_ERR_HIP;
    }
    size_t at = 0;
    auto take = [&](size_t bytes) { size_t here = at; at += align_up(bytes); return here; };
    out.info = take(kInfoWords * sizeof(int));
    out.angular = take((size_t)count * channels * sizeof(double));
    out.records = take(total_nodes * sizeof(NodeRecord));
    out.keys = take(total_nodes * sizeof(int));
    out.ids = take(total_nodes * sizeof(int));
    out.sorted_keys = take(total_nodes * sizeof(int));
    out.sorted_ids = take(total_nodes * sizeof(int));
    out.occupied = take((span + 1) * sizeof(int));
    out.row_scan = take((span + 1) * sizeof(int));
    out.row_key = take(span * sizeof(int));
    out.slice_used = take((size_t)L->positions * sizeof(int));
    out.zero = take((size_t)L->positions * 5 * sizeof(double));
    out.wide = take((size_t)L->positions * sizeof(int));
    out.point_of = take((size_t)L->positions * sizeof(int));
    out.wide_of = take((size_t)L->positions * sizeof(int));
    out.sort_temp = take(sort_temp);
    out.scan_temp = take(scan_temp);
    out.total = at;
    return LH_OK;
}

int make_params(const lh_compile_field* field, const lh_compile_lattice* L, void* workspace,
                size_t bytes, Params& p, Layout& lay) {
    LH_REQUIRE(field != nullptr, "field descriptor is NULL");
    int status = check_lattice(L, field->count);
    if (status != LH_OK) return status;
    LH_REQUIRE(field->start_m && field->direction && field->length_m && field->start_ns
                   && field->end_ns, "field geometry pointers must be set");
    LH_REQUIRE(field->photons ? field->cone_cosine != nullptr
                              : (field->coefficient0 && field->coefficient2 && field->beta),
               "give photons and cone_cosine, or coefficient0, coefficient2 and beta");
    status = make_layout(L, field->count, lay);
    if (status != LH_OK) return status;
    LH_REQUIRE(workspace != nullptr && bytes >= lay.total,
               "workspace too small: %zu bytes given, %zu needed", bytes, lay.total);
    p.field = *field;
    p.lat = *L;
    p.nodes = L->element_order;
    p.channels = channel_count(L->degree, std::min(L->degree, L->azimuthal_degree));
    p.shifts = L->axial_order == 1 ? 2 : 4;
    p.shift_lo = L->axial_order == 1 ? 0 : -1;
    p.ticks = L->time_deposit == 1 ? 2 : 1;
    p.span = (long long)L->kmax * L->positions;
    if (p.nodes == 1) {
        p.node_frac[0] = 0.5;
        p.node_weight[0] = 1.0;
        p.node_frac[1] = p.node_weight[1] = 0.0;
    } else {
        // two-point Gauss-Legendre on [0, 1] (numpy leggauss(2) mapped by (x+1)/2, w/2)
        double x = 1.0 / sqrt(3.0);
        p.node_frac[0] = (-x + 1) / 2;
        p.node_frac[1] = (x + 1) / 2;
        p.node_weight[0] = p.node_weight[1] = 0.5;
    }
    char* base = static_cast<char*>(workspace);
    p.angular = reinterpret_cast<double*>(base + lay.angular);
    p.records = reinterpret_cast<NodeRecord*>(base + lay.records);
    p.keys = reinterpret_cast<int*>(base + lay.keys);
    p.ids = reinterpret_cast<int*>(base + lay.ids);
    p.sorted_keys = reinterpret_cast<int*>(base + lay.sorted_keys);
    p.sorted_ids = reinterpret_cast<int*>(base + lay.sorted_ids);
    p.occupied = reinterpret_cast<int*>(base + lay.occupied);
    p.row_scan = reinterpret_cast<int*>(base + lay.row_scan);
    p.row_key = reinterpret_cast<int*>(base + lay.row_key);
    p.slice_used = reinterpret_cast<int*>(base + lay.slice_used);
    p.zero = reinterpret_cast<double*>(base + lay.zero);
    p.wide = reinterpret_cast<int*>(base + lay.wide);
    p.point_of = reinterpret_cast<int*>(base + lay.point_of);
    p.wide_of = reinterpret_cast<int*>(base + lay.wide_of);
    p.info = reinterpret_cast<int*>(base + lay.info);
    p.points = nullptr;
    return LH_OK;
}

}  // namespace

extern "C" {

size_t lh_compile_workspace_bytes(const lh_compile_lattice* lattice, int count) {
    if (check_lattice(lattice, count) != LH_OK) return 0;
    Layout lay;
    if (make_layout(lattice, count, lay) != LH_OK) return 0;
    return lay.total;
}

int lh_compile_layout(const lh_compile_field* field, const lh_compile_lattice* lattice, int* info,
                      double* points, void* workspace, size_t workspace_bytes, void* stream) {
    Params p;
    Layout lay;
    int status = make_params(field, lattice, workspace, workspace_bytes, p, lay);
    if (status != LH_OK) return status;
    LH_REQUIRE(info != nullptr && points != nullptr, "info and points must be set");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    int total = p.nodes * field->count;
    compile_clear<<<1, kThreads, 0, s>>>(p.info, kInfoWords);
    compile_clear<<<(unsigned)((p.span + kThreads) / kThreads), kThreads, 0, s>>>(p.occupied,
                                                                                 p.span + 1);
    compile_clear<<<(lattice->positions + kThreads - 1) / kThreads, kThreads, 0, s>>>(
        p.slice_used, lattice->positions);
    if ((status = lh::check_launch("compile_clear")) != LH_OK) return status;
    compile_angular<<<(field->count + kThreads - 1) / kThreads, kThreads, 0, s>>>(p);
    if ((status = lh::check_launch("compile_angular")) != LH_OK) return status;
    compile_nodes<<<(total + kThreads - 1) / kThreads, kThreads, 0, s>>>(p);
    if ((status = lh::check_launch("compile_nodes")) != LH_OK) return status;
    size_t sort_bytes = lay.scan_temp - lay.sort_temp;
    // keys are nonnegative: sorted as unsigned on their low bits
    LH_HIP_CALL(lh::sort_pairs(
        base + lay.sort_temp, sort_bytes, reinterpret_cast<unsigned*>(p.keys),
        reinterpret_cast<unsigned*>(p.sorted_keys), p.ids, p.sorted_ids, total, 0,
        bits_for(p.span), s));
    size_t scan_bytes = lay.total - lay.scan_temp;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + lay.scan_temp, scan_bytes, p.occupied,
                                                 p.row_scan, (int)(p.span + 1), s));
    compile_geometry<<<lattice->positions, kThreads, 0, s>>>(p);
    if ((status = lh::check_launch("compile_geometry")) != LH_OK) return status;
    p.points = points;
    compile_layout<<<1, kLayoutThreads, 0, s>>>(p);
    if ((status = lh::check_launch("compile_layout")) != LH_OK) return status;
    long long span = p.span;
    compile_row_keys<<<(unsigned)((span + kThreads - 1) / kThreads), kThreads, 0, s>>>(p);
    if ((status = lh::check_launch("compile_row_keys")) != LH_OK) return status;
    LH_HIP_CALL(hipMemcpyAsync(info, p.info, kInfoWords * sizeof(int), hipMemcpyDeviceToDevice, s));
    return LH_OK;
}

int lh_compile_channels(const lh_compile_field* field, const lh_compile_lattice* lattice,
                        const double* omega, int n_freq, double scale, int rows, double* deposit,
                        void* channels, void* workspace, size_t workspace_bytes, void* stream) {
    Params p;
    Layout lay;
    int status = make_params(field, lattice, workspace, workspace_bytes, p, lay);
    if (status != LH_OK) return status;
    LH_REQUIRE(omega != nullptr && n_freq >= 1 && n_freq <= kMaxFreq,
               "1..%d frequencies are supported", kMaxFreq);
    LH_REQUIRE(rows >= 1 && deposit != nullptr && channels != nullptr,
               "rows, deposit and channels must be set");
    hipStream_t s = static_cast<hipStream_t>(stream);
    {
        lh::KernelMark mark("compile_deposit", s);
        compile_deposit<<<rows, kThreads, deposit_lds_bytes(p.channels), s>>>(
            p, rows, deposit_tile(p.channels), deposit);
    }
    if ((status = lh::check_launch("compile_deposit")) != LH_OK) return status;
    // the phase table follows the rows' deposits in the deposit buffer
    double* table_c = deposit + (size_t)rows * lattice->moments * p.channels;
    double* table_s = table_c + (size_t)rows * n_freq;
    compile_phase_table<<<(unsigned)(((long long)rows * n_freq + kThreads - 1) / kThreads),
                          kThreads, 0, s>>>(p, omega, n_freq, rows, table_c, table_s);
    if ((status = lh::check_launch("compile_phase_table")) != LH_OK) return status;
    // slice and wide counts: the workspace info words of lh_compile_layout
    dim3 grid(lattice->positions, p.channels);
    lh::KernelMark mark("compile_phase", s);
    compile_phase<<<grid, 64, 0, s>>>(p, n_freq, scale, deposit, table_c, table_s,
                                      static_cast<float2*>(channels));
    return lh::check_launch("compile_phase");
}

}  // extern "C"
