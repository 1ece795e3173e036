// Order 1 (single scattering): grouped pencils on gfx950.
//
// Restates the reference's grouped order-1 solver (prompt.py:1020-1258,
// _prompt_numba.py:257-650, 1014-2065, 2247-2354; coarse screen
// _prompt_screen.py:96-477). GPU formulation:
//   compress   one thread per chord sub-segment generates its ring nodes with
//              the reference's azimuth recurrence; nodes are radix-sorted by
//              (cell, pixel) key (stable, so each pencil's nodes stay in
//              enumeration order) and one thread per pencil sums its nodes in
//              that order and finishes the pencil. The screen's parent pencils
//              sum their children's rows in (parent key, child key) order and
//              are laid out in preorder with their children.
//   grouped    one thread per (OM, pencil) computes the pencil's (tau, r,
//              theta) group key (two for the screen's two-node rule); a
//              segmented radix sort per OM orders them by (key, pencil), which
//              is the order in which the reference's hash rows receive their
//              additions; one thread per group sums its members in that order;
//              one workgroup per OM then walks its groups in key order, builds
//              each (tau, r) batch's path grid and profile in parallel over
//              path nodes, and deposits every wavelength's panels with one
//              thread per wavelength (each accumulator has a single writer and
//              a fixed order: deterministic, no atomics).
// All arithmetic is FP64.
#include "lh_common.h"
#include "lh_sort.h"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <cstring>

namespace {

constexpr int kThreads = 256;
constexpr int kNmom = 23;
// Per-sub-segment record of the ring nodes (o1_sub_records), in doubles:
// u e1 e2 p (3 each), cc cs w0 w2 t arc2 sub2, the ring's angle offset, the
// cell index (its bits), one spare.
constexpr int kSubRecord = 22;
constexpr int kMaxMono = 15;        // acceptance degree <= 4
constexpr int kMaxBands = 16;
constexpr int kGauss = 48;
constexpr int kMaxPathNodes = 2048;

__constant__ double kTaylor[3][24];
__constant__ double kTangent[kGauss];
__constant__ double kWeightedSec2[kGauss];

// ------------------------------------------------------------------ compress

__device__ inline long long pixel_of(double sx, double sy, double sz, int n_face) {
    double ax = fabs(sx), ay = fabs(sy), az = fabs(sz), u, v;
    int face;
    if (ax >= ay && ax >= az) { face = sx > 0 ? 0 : 1; u = sy / ax; v = sz / ax; }
    else if (ay >= az) { face = sy > 0 ? 2 : 3; u = sx / ay; v = sz / ay; }
    else { face = sz > 0 ? 4 : 5; u = sx / az; v = sy / az; }
    double quarter = 0.25 * M_PI;
    int i = (int)((atan(u) + quarter) / (2.0 * quarter) * n_face);
    int j = (int)((atan(v) + quarter) / (2.0 * quarter) * n_face);
    i = min(max(i, 0), n_face - 1);
    j = min(max(j, 0), n_face - 1);
    return ((long long)face * n_face + i) * n_face + j;
}

__device__ inline void ring_basis(const double* u, double* e1, double* e2) {
    double h0 = 0, h1 = 0, h2 = 1;
    if (!(fabs(u[2]) < 0.9)) { h0 = 1; h2 = 0; }
    e1[0] = u[1] * h2 - u[2] * h1;
    e1[1] = u[2] * h0 - u[0] * h2;
    e1[2] = u[0] * h1 - u[1] * h0;
    double norm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    e1[0] /= norm; e1[1] /= norm; e1[2] /= norm;
    e2[0] = u[1] * e1[2] - u[2] * e1[1];
    e2[1] = u[2] * e1[0] - u[0] * e1[2];
    e2[2] = u[0] * e1[1] - u[1] * e1[0];
}

struct Compress {
    lh_order1_chords ch;
    lh_order1_grid grid;
    long long nodes;        // subsegments * ring_nodes
    long long* keys;        // (nodes)
    int* ids;               // (nodes)
    long long* sorted_keys;
    int* sorted_ids;
    int key32;              // keys below 2^32: stored and sorted as uint32
    double* sub_record;     // (subsegments, kSubRecord) per-sub-segment moment parts
    int* sub_element;       // (subsegments) unit element index of a sub-segment
    int* sub_index;         // (subsegments) index of the sub-segment in its element
    int* run_flag;          // (nodes + 1)
    int* run_scan;          // (nodes + 1)
    int* run_start;         // (runs + 1)
    long long* run_key;     // (runs)
    double* rows;           // (runs, kNmom)
    int* info;
};

__device__ inline int n_sub(double length, double a_step) {
    int n = (int)ceil(length / a_step);
    return n > 1 ? n : 1;
}

__device__ inline void sub_point(const lh_order1_chords& c, int e, int ia, int n_a, double* p,
                                 double* t) {
    double frac = (ia + 0.5) / n_a;
    for (int k = 0; k < 3; ++k) p[k] = c.start_m[3 * e + k] + frac * c.length_m[e] * c.direction[3 * e + k];
    *t = c.start_ns[e] + frac * (c.end_ns[e] - c.start_ns[e]);
}

__global__ void __launch_bounds__(kThreads) o1_subsegments(Compress k) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.ch.count) return;
    int e = k.ch.index ? k.ch.index[i] : i;
    double a_step = 0.5 * k.grid.cell_m;
    int n_a = n_sub(k.ch.length_m[e], a_step);
    long long first = k.ch.sub_offset[i];
    for (int ia = 0; ia < n_a; ++ia) {
        k.sub_element[first + ia] = i;
        k.sub_index[first + ia] = ia;
    }
}

// Ring nodes of one sub-segment (compress_subsegments, _prompt_numba.py:1122-1190).
// Ring nodes, one thread per node (subsegment s, ring index q): the ring
// direction at angle offset + 2 pi q / R (evaluated directly rather than by
// the reference's rotation recurrence: the same angles to the last bits), its
// cell and pixel key. Consecutive threads write consecutive nodes.
__global__ void __launch_bounds__(kThreads) o1_nodes(Compress k) {
    const int R = k.grid.ring_nodes, face = k.grid.pixel_face;
    long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= k.grid.subsegments * (long long)R) return;
    long long s = node / R;
    int q = (int)(node - s * R);
    // the sub-segment's direction, ring basis, cone, angle offset and cell
    const double* rec = k.sub_record + (size_t)s * kSubRecord;
    double u[3], e1[3], e2[3];
    for (int d = 0; d < 3; ++d) {
        u[d] = rec[d];
        e1[d] = rec[3 + d];
        e2[d] = rec[6 + d];
    }
    const double cc = rec[12], cs = rec[13], offset = rec[19];
    const long long ck = __double_as_longlong(rec[20]);
    long long n_pix = 6LL * face * face;
    double sp, cp;
    sincos(offset + q * (2.0 * M_PI / R), &sp, &cp);
    double sx = cc * u[0] + cs * (cp * e1[0] + sp * e2[0]);
    double sy = cc * u[1] + cs * (cp * e1[1] + sp * e2[1]);
    double sz = cc * u[2] + cs * (cp * e1[2] + sp * e2[2]);
    const long long key = ck * n_pix + pixel_of(sx, sy, sz, face);
    if (k.key32) reinterpret_cast<unsigned*>(k.keys)[node] = (unsigned)key;
    else k.keys[node] = key;
    k.ids[node] = (int)node;
}

// Sorted node key j (stored as uint32 when every key is below 2^32).
__device__ inline long long node_key(const Compress& k, long long j) {
    return k.key32 ? (long long)reinterpret_cast<const unsigned*>(k.sorted_keys)[j]
                   : k.sorted_keys[j];
}

__global__ void __launch_bounds__(kThreads) o1_run_flags(Compress k) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > k.nodes) return;
    if (j == k.nodes) { k.run_flag[j] = 0; return; }
    k.run_flag[j] = (j == 0 || node_key(k, j) != node_key(k, j - 1)) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) o1_run_starts(Compress k) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > k.nodes) return;
    if (j == k.nodes) {
        int runs = k.run_scan[j];
        k.run_start[runs] = (int)k.nodes;
        k.info[0] = runs;
        return;
    }
    if (k.run_flag[j]) {
        int r = k.run_scan[j];
        k.run_start[r] = (int)j;
        k.run_key[r] = node_key(k, j);
    }
}

// Moment row of one (cell, pixel) key: its nodes in enumeration order (:1192-1223).
// Moment contributions of one ring node (the terms summed per pencil by
// compress_ring_pencils, _prompt_numba.py:2246-2279).
// The per-sub-segment part of a ring node's moments: direction, ring basis,
// cone cosine and sine, share weights, sub-point, time and the two spreads.
// Every node of a sub-segment shares it, so it is evaluated once per
// sub-segment (its divisions and square roots were once per node).

__global__ void __launch_bounds__(kThreads) o1_sub_records(Compress k) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= k.grid.subsegments) return;
    const lh_order1_chords& c = k.ch;
    int R = k.grid.ring_nodes;
    double a_step = 0.5 * k.grid.cell_m;
    int i = k.sub_element[s], ia = k.sub_index[s];
    int e = c.index ? c.index[i] : i;
    double u[3] = {c.direction[3 * e], c.direction[3 * e + 1], c.direction[3 * e + 2]};
    double e1[3], e2[3];
    ring_basis(u, e1, e2);
    double cc = c.cone_cosine[e];
    double cs = sqrt(fmax(1.0 - cc * cc, 0.0));
    int n_a = n_sub(c.length_m[e], a_step);
    double share = 1.0 / ((double)n_a * R);
    double arc = 2.0 * M_PI * cs / R;
    double sl = c.length_m[e] / n_a;
    double p[3], t;
    sub_point(c, e, ia, n_a, p, &t);
    double* out = k.sub_record + (size_t)s * kSubRecord;
    for (int d = 0; d < 3; ++d) {
        out[d] = u[d];
        out[3 + d] = e1[d];
        out[6 + d] = e2[d];
        out[9 + d] = p[d];
    }
    out[12] = cc;
    out[13] = cs;
    out[14] = c.coefficient0[e] * share;
    out[15] = c.coefficient2[e] * share;
    out[16] = t;
    out[17] = arc * arc / 12.0;
    out[18] = sl * sl / 12.0;
    const double golden = 0.6180339887498949;
    out[19] = fmod(i * golden, 1.0) * (2.0 * M_PI) / R;
    long long cell[3];
    for (int d = 0; d < 3; ++d) {
        long long v = (long long)floor((p[d] - k.grid.origin[d]) / k.grid.cell_m);
        cell[d] = min(max(v, 0LL), k.grid.dims[d] - 1);
    }
    out[20] = __longlong_as_double((cell[0] * k.grid.dims[1] + cell[1]) * k.grid.dims[2]
                                   + cell[2]);
    out[21] = 0.0;
}

__device__ void node_moments(const Compress& k, int node, double v[kNmom]) {
    int R = k.grid.ring_nodes;
    long long s = node / R;
    const double* rec = k.sub_record + (size_t)s * kSubRecord;
    double u[3], e1[3], e2[3], p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        u[d] = rec[d];
        e1[d] = rec[3 + d];
        e2[d] = rec[6 + d];
        p[d] = rec[9 + d];
    }
    const double cc = rec[12], cs = rec[13], w0 = rec[14], w2 = rec[15], t = rec[16];
    const double arc2 = rec[17], sub2 = rec[18];
    // the ring direction's angle, evaluated as o1_nodes evaluates it
    const int q = (int)(node - s * R);
    const double offset = rec[19];
    double sp, cp;
    sincos(offset + q * (2.0 * M_PI / R), &sp, &cp);
    double sx = cc * u[0] + cs * (cp * e1[0] + sp * e2[0]);
    double sy = cc * u[1] + cs * (cp * e1[1] + sp * e2[1]);
    double sz = cc * u[2] + cs * (cp * e1[2] + sp * e2[2]);
    double px = p[0], py = p[1], pz = p[2];
    v[0] = w0;
    v[1] = w2;
    v[2] = w0 * px;
    v[3] = w0 * py;
    v[4] = w0 * pz;
    v[5] = w0 * t;
    v[6] = w0 * sx;
    v[7] = w0 * sy;
    v[8] = w0 * sz;
    v[9] = w0 * (px * px + py * py + pz * pz + sub2);
    v[10] = w0 * t * t;
    double tx = -sp * e1[0] + cp * e2[0];
    double ty = -sp * e1[1] + cp * e2[1];
    double tz = -sp * e1[2] + cp * e2[2];
    v[11] = w0 * (sx * sx + arc2 * tx * tx);
    v[12] = w0 * (sy * sy + arc2 * ty * ty);
    v[13] = w0 * (sz * sz + arc2 * tz * tz);
    v[14] = w0 * (sx * sy + arc2 * tx * ty);
    v[15] = w0 * (sx * sz + arc2 * tx * tz);
    v[16] = w0 * (sy * sz + arc2 * ty * tz);
    v[17] = w0 * (px * px + sub2 * u[0] * u[0]);
    v[18] = w0 * (py * py + sub2 * u[1] * u[1]);
    v[19] = w0 * (pz * pz + sub2 * u[2] * u[2]);
    v[20] = w0 * (px * py + sub2 * u[0] * u[1]);
    v[21] = w0 * (px * pz + sub2 * u[0] * u[2]);
    v[22] = w0 * (py * pz + sub2 * u[1] * u[2]);
}

// Wave sums of V <= 32 values per lane, reduce-scattered: every step pairs
// lanes `o` apart, each keeps one half of its values and adds the partner's
// copy of that half (32 values padded, offsets 32 .. 2: 16 + 8 + 4 + 2 + 1
// shuffles, then one across lanes 1 apart). Lane l ends with the wave total of
// value l >> 1 (both lanes of a pair hold it). A fixed order, as a butterfly
// per value would be, with a fifth of its shuffles.
template <int V>
__device__ inline double wave_totals(const double (&v)[V], int lane) {
    static_assert(V >= 1 && V <= 32, "at most 32 values");
    double a[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) a[i] = i < V ? v[i] : 0.0;
#pragma unroll
    for (int step = 0; step < 5; ++step) {
        const int o = 32 >> step, h = 16 >> step;
        const bool upper = (lane & o) != 0;
#pragma unroll
        for (int i = 0; i < h; ++i) {
            double send = upper ? a[i] : a[h + i];
            double keep = upper ? a[h + i] : a[i];
            a[i] = keep + __shfl_xor(send, o, 64);
        }
    }
    return a[0] + __shfl_xor(a[0], 1, 64);
}

// Block b of the first `used` blocks (dispatched to XCD b % 8 in turn)
// renumbered so that every XCD's blocks cover one contiguous range.
__device__ inline long long xcd_block(long long b, long long used) {
    constexpr int kXcds = 8;
    const long long per = used / kXcds, rem = used % kXcds;
    const long long xcd = b % kXcds, k = b / kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

constexpr int kRowPiece = 64;   // ring nodes per wave of the moment rows

// Pieces of every run: ceil(nodes / kRowPiece); runs of more than one piece
// also count their pieces in `big` (their partial rows are kept).
__global__ void __launch_bounds__(kThreads) o1_row_counts(Compress k, int runs, int* counts,
                                                          int* big) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > runs) return;
    int c = r == runs ? 0 : (k.run_start[r + 1] - k.run_start[r] + kRowPiece - 1) / kRowPiece;
    counts[r] = c;
    big[r] = c > 1 ? c : 0;
}

// The run of every piece, one thread per run writing its pieces.
__global__ void __launch_bounds__(kThreads) o1_piece_runs(int runs, const int* piece_start,
                                                          int* piece_run) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    for (int q = piece_start[r]; q < piece_start[r + 1]; ++q) piece_run[q] = r;
}

// Moment rows: one wave per piece of at most kRowPiece nodes of one run, a
// fixed butterfly sum; a run of one piece writes its row, longer runs write
// partial rows that o1_row_combine adds in piece order (deterministic).
__global__ void __launch_bounds__(kThreads) o1_row_pieces(Compress k, int runs,
                                                          const int* piece_start,
                                                          const int* big_start,
                                                          const int* piece_run,
                                                          double* piece_rows) {
    int lane = threadIdx.x & 63;
    // each XCD takes a contiguous range of pieces (neighbouring cells share
    // elements), as in o1_group_sums
    const long long used = ((long long)piece_start[runs] + kThreads / 64 - 1) / (kThreads / 64);
    if ((long long)blockIdx.x >= used) return;
    int q = (int)(xcd_block(blockIdx.x, used) * (kThreads / 64)) + (threadIdx.x >> 6);
    if (q >= piece_start[runs]) return;
    int r = piece_run[q], piece = q - piece_start[r];
    int j0 = k.run_start[r] + piece * kRowPiece;
    int j1 = min(k.run_start[r + 1], j0 + kRowPiece);
    double v[kNmom];
    if (j0 + lane < j1) {
        node_moments(k, k.sorted_ids[j0 + lane], v);
    } else {
        for (int m = 0; m < kNmom; ++m) v[m] = 0.0;
    }
    bool single = piece_start[r + 1] - piece_start[r] == 1;
    double* out = single ? k.rows + (size_t)r * kNmom
                         : piece_rows + ((size_t)big_start[r] + piece) * kNmom;
    double total = wave_totals<kNmom>(v, lane);
    if ((lane & 1) == 0 && (lane >> 1) < kNmom) out[lane >> 1] = total;
}

// Partial rows of a run added in piece order, one thread per run (most runs
// are one piece, written by o1_row_pieces, and return at once).
__global__ void __launch_bounds__(kThreads) o1_row_combine(Compress k, int runs,
                                                           const int* piece_start,
                                                           const int* big_start,
                                                           const double* piece_rows) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    int n = piece_start[r + 1] - piece_start[r];
    if (n < 2) return;
    const double* in = piece_rows + (size_t)big_start[r] * kNmom;
    double total[kNmom];
#pragma unroll
    for (int m = 0; m < kNmom; ++m) total[m] = in[m];
    for (int i = 1; i < n; ++i) {
#pragma unroll
        for (int m = 0; m < kNmom; ++m) total[m] += in[(size_t)i * kNmom + m];
    }
#pragma unroll
    for (int m = 0; m < kNmom; ++m) k.rows[(size_t)r * kNmom + m] = total[m];
}

// Pencil from a moment row (finish_pencils, _prompt_numba.py:2282-2353).
__device__ void finish_row(const double* a, const lh_order1_pencils& out, int q) {
    double w = a[0];
    double mx = a[2] / w, my = a[3] / w, mz = a[4] / w;
    out.position[3 * q] = mx; out.position[3 * q + 1] = my; out.position[3 * q + 2] = mz;
    out.time[q] = a[5] / w;
    out.w0[q] = w;
    out.w2[q] = a[1];
    double rx = a[6] / w, ry = a[7] / w, rz = a[8] / w;
    double norm = sqrt(rx * rx + ry * ry + rz * rz);
    double d[3] = {rx / norm, ry / norm, rz / norm};
    out.direction[3 * q] = d[0]; out.direction[3 * q + 1] = d[1]; out.direction[3 * q + 2] = d[2];
    double cxx = a[11] / w - rx * rx, cyy = a[12] / w - ry * ry, czz = a[13] / w - rz * rz;
    double cxy = a[14] / w - rx * ry, cxz = a[15] / w - rx * rz, cyz = a[16] / w - ry * rz;
    double e1[3], e2[3];
    ring_basis(d, e1, e2);
    double c11 = (e1[0] * (cxx * e1[0] + cxy * e1[1] + cxz * e1[2])
                  + e1[1] * (cxy * e1[0] + cyy * e1[1] + cyz * e1[2])
                  + e1[2] * (cxz * e1[0] + cyz * e1[1] + czz * e1[2]));
    double c22 = (e2[0] * (cxx * e2[0] + cxy * e2[1] + cxz * e2[2])
                  + e2[1] * (cxy * e2[0] + cyy * e2[1] + cyz * e2[2])
                  + e2[2] * (cxz * e2[0] + cyz * e2[1] + czz * e2[2]));
    double c12 = (e1[0] * (cxx * e2[0] + cxy * e2[1] + cxz * e2[2])
                  + e1[1] * (cxy * e2[0] + cyy * e2[1] + cyz * e2[2])
                  + e1[2] * (cxz * e2[0] + cyz * e2[1] + czz * e2[2]));
    double half = 0.5 * (c11 + c22);
    double disc = sqrt(fmax(0.25 * (c11 - c22) * (c11 - c22) + c12 * c12, 0.0));
    out.var_major[q] = fmax(half + disc, 0.0);
    out.var_minor[q] = fmax(half - disc, 0.0);
    double angle = 0.5 * atan2(2.0 * c12, c11 - c22);
    double ca = cos(angle), sa = sin(angle);
    for (int k = 0; k < 3; ++k) out.major[3 * q + k] = ca * e1[k] + sa * e2[k];
    out.pos_cov[6 * q] = fmax(a[17] / w - mx * mx, 0.0);
    out.pos_cov[6 * q + 1] = fmax(a[18] / w - my * my, 0.0);
    out.pos_cov[6 * q + 2] = fmax(a[19] / w - mz * mz, 0.0);
    out.pos_cov[6 * q + 3] = a[20] / w - mx * my;
    out.pos_cov[6 * q + 4] = a[21] / w - mx * mz;
    out.pos_cov[6 * q + 5] = a[22] / w - my * mz;
}

__global__ void __launch_bounds__(kThreads) o1_finish(const double* rows, int runs,
                                                      lh_order1_pencils out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    finish_row(rows + (size_t)r * kNmom, out, r);
}

// Screen hierarchy: parent key of a fine (cell, pixel) key (_prompt_screen.py:44-53).
__device__ inline long long parent_key(long long key, int face, int parent_face) {
    int factor = face / parent_face;
    long long per_cell = 6LL * face * face;
    long long cell = key / per_cell, pixel = key % per_cell;
    long long cube = pixel / ((long long)face * face), ij = pixel % ((long long)face * face);
    long long i = ij / face, j = ij % face;
    return cell * (6LL * parent_face * parent_face)
           + (cube * parent_face + i / factor) * parent_face + j / factor;
}

__global__ void __launch_bounds__(kThreads) o1_parent_keys(Compress k, int runs, long long* pkeys,
                                                           int* pids) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    pkeys[r] = parent_key(k.run_key[r], k.grid.pixel_face, k.grid.parent_face);
    pids[r] = r;
}

// After sorting fine pencils by (parent key, fine key): parent runs.
__global__ void __launch_bounds__(kThreads) o1_parent_flags(const long long* spkeys, int runs,
                                                            int* flag) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > runs) return;
    flag[j] = j == runs ? 0 : ((j == 0 || spkeys[j] != spkeys[j - 1]) ? 1 : 0);
}

// Parent rows (children's rows summed in fine-key order), preorder layout:
// parent c at position start_c + c, its children right after it.
__global__ void __launch_bounds__(kThreads) o1_parents(const double* rows, const int* spids,
                                                       const int* flag, const int* scan,
                                                       int runs, lh_order1_pencils out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= runs) return;
    int c = scan[j] + (flag[j] ? 0 : -1);     // parent index of sorted child j
    int position = j + c + 1;                  // preorder slot of the child
    finish_row(rows + (size_t)spids[j] * kNmom, out, position);
    if (!flag[j]) return;
    // first child of parent c: sum the parent's rows and write the parent
    int end = j + 1;
    while (end < runs && !flag[end]) ++end;
    double acc[kNmom];
    for (int m = 0; m < kNmom; ++m) acc[m] = 0.0;
    for (int q = j; q < end; ++q)
        for (int m = 0; m < kNmom; ++m) acc[m] += rows[(size_t)spids[q] * kNmom + m];
    int slot = j + c;
    finish_row(acc, out, slot);
    out.parent[slot] = -2;
    out.next[slot] = end + c + 1;
}

__global__ void __launch_bounds__(kThreads) o1_parent_links(const int* flag, const int* scan,
                                                            int runs, lh_order1_pencils out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= runs) return;
    int c = scan[j] + (flag[j] ? 0 : -1);
    int position = j + c + 1;
    // the parent's slot is the first child's position minus one
    int first = j;
    while (!flag[first]) --first;
    out.parent[position] = first + c;
    out.next[position] = position + 1;
}

// ------------------------------------------------------------------- grouped

struct BatchHead;

struct Group {
    lh_order1_pencils pen;
    lh_order1_oms om;
    lh_order1_optics opt;
    lh_order1_merge merge;
    int compute_bins;
    int screen;
    int n_pen;
    int slots;                // entries per (pencil, OM): 2 on the screen, else 1
    int key_bits;             // bits of the compact per-OM key; the OM sits above them
    long long key_mask;       // (1 << key_bits) - 1, also the compact sentinel
    int key32;                // the keys fit 32 bits: stored and sorted as uint32
    long long theta_bins;
    double log_ratio;
    double mut_min, speed_max;
    long long* keys;          // (work, slots * n_pen) (OM << key_bits) | compact key
    int* vals;
    long long* sorted_keys;
    int* sorted_vals;
    int* group_flag;          // (work * slots * n_pen + 1)
    int* group_scan;
    int* group_start;         // (groups + 1)
    int* group_end;           // (groups + 1) one past the last member of every group
    int* member_off;          // (groups + 1) first member of every group, numbered
                              // consecutively over the groups           caller scratch
    int* piece_start;         // (groups + 1) first piece of every group   caller scratch
    int* tile_group;          // (tiles) group of every tile's first member  caller scratch
    double* piece_values;     // (pieces, nv) partial group sums          caller scratch
    int* om_groups;           // (work + 1) first group of every OM
    double* values;           // (groups, nv)            caller scratch
    int* batch_flag;          // (groups + 1)            caller scratch
    int* batch_scan;          // (groups + 1)
    int* batch_start;         // (batches + 1) first group of every batch
    int* om_batches;          // (work + 1) first batch of every OM
    double* part_acc;         // (work, segments, bands, bins, 4) partial deposits
    double* part_pre;         // (work, segments, bands, 3)
    double* part_charge;      // (work, segments, bands)
    int segments;             // batch segments per OM
    BatchHead* heads;         // (groups + 1) batch headers           caller scratch
    long long* node_n;        // (groups + 1) path nodes per batch
    long long* node_off;      // (groups + 1) first node of every batch in the pool
    double* pool_path;        // (pool_nodes) path nodes of every batch
    double* pool_g0;          // (pool_nodes) profile sums
    double* pool_g2;
    long long pool_nodes;     // nodes the pool holds (0: the single-kernel form)
    unsigned long long* work_counts;  // optional: profile terms, panels, bin events,
                                      // group members
    double* prepared;         // (n_pen, 7)
    double* lambda_bins;      // (work, bands, bins)
    int* k_range;             // (work, 4) k1 min, k1 max, k0 min, k0 max
    unsigned long long* x_range;  // (work, 4) ordered r min, r max, tau min, tau max
    unsigned long long* max_key;
    double* charge;           // (n_recv) output
    double* bins;             // (n_recv, bins) output
    int* info;
};

__global__ void __launch_bounds__(kThreads) o1_prepare_shape(Group G) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.n_pen) return;
    const double* a = G.pen.major + 3 * p;
    const double* s = G.pen.direction + 3 * p;
    const double* cov = G.pen.pos_cov + 6 * p;
    double bx = s[1] * a[2] - s[2] * a[1];
    double by = s[2] * a[0] - s[0] * a[2];
    double bz = s[0] * a[1] - s[1] * a[0];
    double* q = G.prepared + 7 * (size_t)p;
    q[0] = bx; q[1] = by; q[2] = bz;
    q[3] = (a[0] * (cov[0] * a[0] + cov[3] * a[1] + cov[4] * a[2])
            + a[1] * (cov[3] * a[0] + cov[1] * a[1] + cov[5] * a[2])
            + a[2] * (cov[4] * a[0] + cov[5] * a[1] + cov[2] * a[2]));
    q[4] = (bx * (cov[0] * bx + cov[3] * by + cov[4] * bz)
            + by * (cov[3] * bx + cov[1] * by + cov[5] * bz)
            + bz * (cov[4] * bx + cov[5] * by + cov[2] * bz));
    q[5] = (a[0] * (cov[0] * bx + cov[3] * by + cov[4] * bz)
            + a[1] * (cov[3] * bx + cov[1] * by + cov[5] * bz)
            + a[2] * (cov[4] * bx + cov[5] * by + cov[2] * bz));
    q[6] = cov[0] + cov[1] + cov[2];
}

__device__ double gauss_inverse_mean(double u, double v, double var_u, double var_v) {
    double scale = 2.0 * (var_u + var_v) + u * u + v * v;
    if (scale <= 0.0) return 1e300;
    double s0 = 1.0 / sqrt(scale);
    double total = 0.0;
#pragma unroll 2
    for (int k = 0; k < kGauss; ++k) {
        double s = s0 * kTangent[k];
        double s2 = s * s;
        double a = 1.0 + 2.0 * s2 * var_u;
        double b = 1.0 + 2.0 * s2 * var_v;
        total += kWeightedSec2[k] * exp(-s2 * (u * u / a + v * v / b)) / sqrt(a * b);
    }
    return 2.0 / sqrt(M_PI) * s0 * total;
}

// _pencil_shape_factor_from_ray (_prompt_numba.py:1389-1436).
__device__ bool shape_factor(const Group& G, int p, double r, double theta, double sine,
                             double hx, double hy, double hz, double* factor_out,
                             double* theta_out, int* averaged_out) {
    const double* a = G.pen.major + 3 * p;
    const double* q = G.prepared + 7 * (size_t)p;
    double smoothing = G.opt.smoothing;
    double inv_r2 = 1.0 / (r * r);
    double m11 = G.pen.var_major[p] + q[3] * inv_r2;
    double m22 = G.pen.var_minor[p] + q[4] * inv_r2;
    double m12 = q[5] * inv_r2;
    double upper = fmax(m11, m22) + fabs(m12);
    double biggest = fmax(fmax(fabs(m11), fabs(m22)), fabs(m12));
    bool far = biggest < 1e150
               && theta * theta > smoothing * smoothing * fmax(upper, 0.0) * (1.0 + 1e-12);
    double mu1 = 0, mu2 = 0, extent = 0;
    if (!far) {
        double half_tr = 0.5 * (m11 + m22);
        double disc = sqrt(fmax(0.25 * (m11 - m22) * (m11 - m22) + m12 * m12, 0.0));
        mu1 = half_tr + disc;
        mu2 = fmax(half_tr - disc, 1e-16 * mu1 + 1e-300);
        extent = sqrt(fmax(mu1, 0.0));
    }
    double u = hx * a[0] + hy * a[1] + hz * a[2];
    double v = hx * q[0] + hy * q[1] + hz * q[2];
    if (!far && extent > 0.0 && theta < smoothing * extent) {
        double angle = 0.5 * atan2(2.0 * m12, m11 - m22);
        double ca = cos(angle), sa = sin(angle);
        double mean = gauss_inverse_mean(ca * u + sa * v, -sa * u + ca * v, mu1, mu2);
        double factor = mean * (sine > 1e-300 ? theta / sine : 1.0);
        *theta_out = factor > 1.0 ? asin(fmin(1.0, 1.0 / factor)) : theta;
        *factor_out = factor;
        *averaged_out = theta < extent;
        return true;
    }
    if (sine <= 1e-300) { *factor_out = 0.0; *theta_out = theta; *averaged_out = 0; return false; }
    double d2 = u * u + v * v;
    double factor = 1.0 / sine;
    if (d2 > 0.0) {
        double quad_form = m11 * u * u + 2.0 * m12 * u * v + m22 * v * v;
        factor *= 1.0 + 0.5 * (3.0 * quad_form / d2 - (m11 + m22)) / d2;
    }
    *factor_out = factor;
    *theta_out = theta;
    *averaged_out = 0;
    return true;
}

// inverse_radius_rule + pencil_radial_rule (_prompt_screen.py:96-167).
__device__ bool radial_rule(const Group& G, int p, double r, double theta, double sine, double hx,
                            double hy, double hz, double* f1, double* t1, double* f2, double* t2,
                            int* averaged) {
    const double* q = G.prepared + 7 * (size_t)p;
    const double* a = G.pen.major + 3 * p;
    double smoothing = G.opt.smoothing;
    double m11 = G.pen.var_major[p] + q[3] / (r * r);
    double m22 = G.pen.var_minor[p] + q[4] / (r * r);
    double m12 = q[5] / (r * r);
    double half = 0.5 * (m11 + m22);
    double disc = sqrt(fmax(0.25 * (m11 - m22) * (m11 - m22) + m12 * m12, 0.0));
    double mu1 = half + disc;
    double mu2 = fmax(half - disc, 1e-16 * mu1 + 1e-300);
    double extent = sqrt(fmax(mu1, 0.0));
    if (extent > 0.0 && theta < smoothing * extent) {
        double u = hx * a[0] + hy * a[1] + hz * a[2];
        double v = hx * q[0] + hy * q[1] + hz * q[2];
        if (sine > 0.0) { u *= theta / sine; v *= theta / sine; }
        double angle = 0.5 * atan2(2 * m12, m11 - m22);
        double ca = cos(angle), sa = sin(angle);
        double uu = ca * u + sa * v, vv = -sa * u + ca * v;
        double scale = 2.0 * (mu1 + mu2) + uu * uu + vv * vv;
        if (scale > 0.0) {
            double s0 = 1.0 / sqrt(scale);
            double inverse = 0.0, radial = 0.0;
            for (int k = 0; k < kGauss; ++k) {
                double t = s0 * kTangent[k];
                double aa = 1.0 + 2.0 * t * t * mu1;
                double bb = 1.0 + 2.0 * t * t * mu2;
                double mass = kWeightedSec2[k] * exp(-t * t * (uu * uu / aa + vv * vv / bb))
                              / sqrt(aa * bb);
                inverse += mass;
                radial += mass * (mu1 / aa + mu2 / bb + uu * uu / (aa * aa) + vv * vv / (bb * bb));
            }
            double multiplier = 2.0 / sqrt(M_PI) * s0;
            inverse *= multiplier;
            radial *= multiplier;
            double w0 = 0, a0 = 0, w1 = 0, a1 = 0;
            if (inverse > 0.0) {
                double mean = 1.0 / inverse;
                double second = radial / inverse;
                double variance = second - mean * mean;
                if (variance <= 0.0) {
                    w0 = inverse; a0 = mean; w1 = 0.0; a1 = mean;
                } else {
                    double third = (uu * uu + vv * vv + mu1 + mu2) / inverse;
                    double other = (third - 2.0 * mean * second + mean * mean * mean) / variance;
                    double dd = sqrt((mean - other) * (mean - other) + 4.0 * variance);
                    double low = 0.5 * (mean + other - dd);
                    double high = 0.5 * (mean + other + dd);
                    double low_weight = inverse * (high - mean) / (high - low);
                    w0 = low_weight; a0 = low; w1 = inverse - low_weight; a1 = high;
                }
            }
            if (a0 > 0.0 && a1 < 0.5 * M_PI && w0 > 0.0 && w1 >= 0.0) {
                *f1 = w0 * a0 / sin(a0); *t1 = a0;
                *f2 = w1 * a1 / sin(a1); *t2 = a1;
                *averaged = theta < extent;
                return true;
            }
        }
    }
    double factor, angle;
    bool ok = shape_factor(G, p, r, theta, sine, hx, hy, hz, &factor, &angle, averaged);
    *f1 = factor; *t1 = angle; *f2 = 0.0; *t2 = angle;
    return ok;
}

struct Entry {
    long long key;
    double w0, w2, c1, c3, theta, r, t_emit, tau;
    int averaged;
};

__device__ inline int theta_index(double theta, double theta_lo, double log_ratio) {
    if (theta < theta_lo) return 0;
    return 1 + (int)(log(theta / theta_lo) / log_ratio);
}

__device__ inline void frame_of(const Group& G, int p, int d, double* r, double* h, double* theta,
                                double* sine, double* c1, double* c3) {
    const double* pp = G.pen.position + 3 * p;
    double dx = G.om.positions[3 * d] - pp[0];
    double dy = G.om.positions[3 * d + 1] - pp[1];
    double dz = G.om.positions[3 * d + 2] - pp[2];
    double rr = sqrt(dx * dx + dy * dy + dz * dz);
    h[0] = dx / rr; h[1] = dy / rr; h[2] = dz / rr;
    const double* s = G.pen.direction + 3 * p;
    double nx = G.om.looks[3 * d], ny = G.om.looks[3 * d + 1], nz = G.om.looks[3 * d + 2];
    *c1 = s[0] * nx + s[1] * ny + s[2] * nz;
    double cosine = h[0] * s[0] + h[1] * s[1] + h[2] * s[2];
    double px = h[0] - cosine * s[0], py = h[1] - cosine * s[1], pz = h[2] - cosine * s[2];
    double sn = sqrt(px * px + py * py + pz * pz);
    *c3 = sn > 1e-300 ? (px * nx + py * ny + pz * nz) / sn : 0.0;
    *theta = atan2(sn, cosine);
    *sine = sn;
    *r = rr;
}

// The (up to two) group entries of pencil p at OM d (_grouped_pencils_order1_block
// :1815-1922; screen :222-326).
__device__ inline void radius_keys(const Group& G, int p, int d, long long* k0, long long* k1) {
    const double* pp = G.pen.position + 3 * p;
    double dx = G.om.positions[3 * d] - pp[0];
    double dy = G.om.positions[3 * d + 1] - pp[1];
    double dz = G.om.positions[3 * d + 2] - pp[2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    *k0 = (long long)floor(log(r) / G.merge.log_r_step) + (1LL << 20);
    if (G.screen) {
        *k1 = 0;
        return;
    }
    double t_emit = G.pen.time[p] - G.om.origins[d];
    double tau = t_emit + r / G.merge.speed_ref;
    long long k = (long long)floor(tau / G.merge.tau_step) + (1LL << 20);
    *k1 = min(max(k, 0LL), (1LL << 21) - 1);
}

// Keys are compact per OM: ((k1 - k1min) * n_k0 + (k0 - k0min)) * 2 theta_bins
// + backward * theta_bins + theta index. They order and group exactly as the
// reference's (k1 * 2^21 + k0) * 2 theta_bins + ... keys.
// Fills `out` with entry `want` (0 or 1) and, when key_other is set and there
// are two entries, the key of entry 1; returns the number of entries.
__device__ int pencil_entries(const Group& G, int p, int d, int w, int want, Entry& out,
                              long long* key_other) {
    double r, h[3], theta, sine, c1, c3;
    frame_of(G, p, d, &r, h, &theta, &sine, &c1, &c3);
    if (G.screen) {
        long long parent = G.pen.parent[p];
        if (parent == -2) {
            double extent2 = G.pen.var_major[p] + G.prepared[7 * (size_t)p + 6] / (r * r);
            if (theta < 0.5 * M_PI && theta * theta < G.opt.smoothing * G.opt.smoothing * extent2)
                return 0;
        } else {
            int q = (int)parent;
            double pr, ph[3], pt, ps, pc1, pc3;
            frame_of(G, q, d, &pr, ph, &pt, &ps, &pc1, &pc3);
            double extent2 = G.pen.var_major[q] + G.prepared[7 * (size_t)q + 6] / (pr * pr);
            bool near = pt < 0.5 * M_PI && pt * pt < G.opt.smoothing * G.opt.smoothing * extent2;
            if (!near) return 0;   // the parent stands for its children
        }
    }
    double t_emit = G.screen ? 0.0 : G.pen.time[p] - G.om.origins[d];
    double tau = G.screen ? 0.0 : t_emit + r / G.merge.speed_ref;
    long long base;
    {
        long long k0 = (long long)floor(log(r) / G.merge.log_r_step) + (1LL << 20);
        long long k1 = 0;
        if (!G.screen) {
            k1 = (long long)floor(tau / G.merge.tau_step) + (1LL << 20);
            k1 = min(max(k1, 0LL), (1LL << 21) - 1);
        }
        const int* range = G.k_range + 4 * w;
        long long n_k0 = (long long)range[3] - range[2] + 1;
        long long n_k1 = (long long)range[1] - range[0] + 1;
        long long d1 = min(max(k1 - range[0], 0LL), n_k1 - 1);
        long long d0 = min(max(k0 - range[2], 0LL), n_k0 - 1);
        base = (d1 * n_k0 + d0) * 2 * G.theta_bins;
    }
    int backward = theta > 0.5 * M_PI;
    double bound = 0.0;
    if (!backward)
        bound = G.pen.var_major[p] + G.pen.var_minor[p] + G.prepared[7 * (size_t)p + 6] / (r * r);
    double weight, weight2 = 0.0, theta2 = theta;
    int averaged;
    if (backward) {
        weight = 1.0 / r;
        averaged = 0;
    } else if (sine > 1e-300 && bound < 1e-4 * sine * sine) {
        weight = (1.0 / sine) / r;
        averaged = 0;
    } else if (G.screen) {
        double f1, t1, f2, t2;
        if (!radial_rule(G, p, r, theta, sine, h[0], h[1], h[2], &f1, &t1, &f2, &t2, &averaged))
            return 0;
        weight = f1 / r;
        weight2 = f2 / r;
        theta = t1;
        theta2 = t2;
    } else {
        double factor, theta_eff;
        if (!shape_factor(G, p, r, theta, sine, h[0], h[1], h[2], &factor, &theta_eff, &averaged))
            return 0;
        weight = factor / r;
        theta = theta_eff;
    }
    int nodes = weight2 > 0.0 ? 2 : 1;
    double wsel = want == 0 ? weight : weight2;
    double th = want == 0 ? theta : theta2;
    out.key = base + backward * G.theta_bins + theta_index(th, G.merge.theta_lo, G.log_ratio);
    out.w0 = wsel * G.pen.w0[p];
    out.w2 = wsel * G.pen.w2[p];
    out.c1 = c1;
    out.c3 = c3;
    out.averaged = averaged;
    out.theta = th;
    out.r = r;
    out.t_emit = t_emit;
    out.tau = tau;
    if (key_other != nullptr && nodes == 2)
        *key_other = base + backward * G.theta_bins
                     + theta_index(theta2, G.merge.theta_lo, G.log_ratio);
    return nodes;
}

__global__ void __launch_bounds__(kThreads) o1_range_init(Group G, int work) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w == 0) G.max_key[0] = 0ULL;
    if (w >= work) return;
    G.x_range[4 * w] = ~0ULL;
    G.x_range[4 * w + 1] = 0ULL;
    G.x_range[4 * w + 2] = ~0ULL;
    G.x_range[4 * w + 3] = 0ULL;
}


// Orderable 64-bit encoding of a double (unsigned order = numeric order).
__device__ inline unsigned long long order_bits(double x) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

__device__ inline double from_order_bits(unsigned long long b) {
    return __longlong_as_double((long long)((b >> 63) ? (b & 0x7fffffffffffffffULL) : ~b));
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Radius and delay extremes of every OM over its pencils: each thread strides
// over pencils, the workgroup reduces in LDS, and one thread adds the result
// with atomics on the ordered encodings (a handful of workgroups per OM keeps
// the atomics few). The keys floor(log r / step) and floor(tau / step) do not
// decrease with r and tau, so their ranges follow from the extremes.
constexpr int kRangeGroups = 16;   // workgroups per OM

__global__ void __launch_bounds__(kThreads) o1_group_ranges(Group G) {
    __shared__ double s_lo[kThreads / 64][2], s_hi[kThreads / 64][2];
    int w = blockIdx.y;
    int d = G.om.work[w];
    const double inf = __builtin_huge_val();
    double r_lo = inf, r_hi = -inf, t_lo = inf, t_hi = -inf;
    const double ox = G.om.positions[3 * d], oy = G.om.positions[3 * d + 1],
                 oz = G.om.positions[3 * d + 2], origin = G.om.origins[d];
    for (int p = blockIdx.x * kThreads + threadIdx.x; p < G.n_pen;
         p += gridDim.x * kThreads) {
        const double* pp = G.pen.position + 3 * p;
        double dx = ox - pp[0], dy = oy - pp[1], dz = oz - pp[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        r_lo = fmin(r_lo, r);
        r_hi = fmax(r_hi, r);
        if (!G.screen) {
            double tau = (G.pen.time[p] - origin) + r / G.merge.speed_ref;
            t_lo = fmin(t_lo, tau);
            t_hi = fmax(t_hi, tau);
        }
    }
    r_lo = wave_min(r_lo);
    r_hi = wave_max(r_hi);
    t_lo = wave_min(t_lo);
    t_hi = wave_max(t_hi);
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_lo[wave][0] = r_lo; s_hi[wave][0] = r_hi;
        s_lo[wave][1] = t_lo; s_hi[wave][1] = t_hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; ++k) {
            r_lo = fmin(r_lo, s_lo[k][0]); r_hi = fmax(r_hi, s_hi[k][0]);
            t_lo = fmin(t_lo, s_lo[k][1]); t_hi = fmax(t_hi, s_hi[k][1]);
        }
        if (r_lo <= r_hi) {
            unsigned long long* x = G.x_range + 4 * w;
            atomicMin(x, order_bits(r_lo));
            atomicMax(x + 1, order_bits(r_hi));
            if (!G.screen) {
                atomicMin(x + 2, order_bits(t_lo));
                atomicMax(x + 3, order_bits(t_hi));
            }
        }
    }
}

// Key ranges of every OM (one key of margin on each side: the key pass
// evaluates the same expressions elsewhere, where contraction may differ in the
// last bit) and the largest compact key space.
__global__ void __launch_bounds__(kThreads) o1_key_space(Group G, int work) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= work) return;
    const unsigned long long* x = G.x_range + 4 * w;
    long long k0_lo = (long long)floor(log(from_order_bits(x[0])) / G.merge.log_r_step)
                      + (1LL << 20);
    long long k0_hi = (long long)floor(log(from_order_bits(x[1])) / G.merge.log_r_step)
                      + (1LL << 20);
    long long k1_lo = 0, k1_hi = 0;
    if (!G.screen) {
        k1_lo = (long long)floor(from_order_bits(x[2]) / G.merge.tau_step) + (1LL << 20);
        k1_hi = (long long)floor(from_order_bits(x[3]) / G.merge.tau_step) + (1LL << 20);
        k1_lo = min(max(k1_lo, 0LL), (1LL << 21) - 1);
        k1_hi = min(max(k1_hi, 0LL), (1LL << 21) - 1);
    }
    int* range = G.k_range + 4 * w;
    range[0] = (int)k1_lo - 1;
    range[1] = (int)k1_hi + 1;
    range[2] = (int)k0_lo - 1;
    range[3] = (int)k0_hi + 1;
    unsigned long long n_k1 = (unsigned long long)(k1_hi - k1_lo + 3);
    unsigned long long n_k0 = (unsigned long long)(k0_hi - k0_lo + 3);
    atomicMax(G.max_key, n_k1 * n_k0 * 2ULL * (unsigned long long)G.theta_bins);
}

__global__ void __launch_bounds__(kThreads) o1_group_keys(Group G) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int w = blockIdx.y;
    if (p >= G.n_pen) return;
    int d = G.om.work[w];
    Entry e0;
    long long key1 = 0;
    int n = pencil_entries(G, p, d, w, 0, e0, &key1);
    size_t at = ((size_t)w * G.n_pen + p) * G.slots;
    long long high = (long long)w << G.key_bits;
    const long long k0 = high | (n > 0 ? e0.key : G.key_mask);
    const long long k1 = high | (n > 1 ? key1 : G.key_mask);
    if (G.key32) {
        unsigned* keys = reinterpret_cast<unsigned*>(G.keys);
        keys[at] = (unsigned)k0;
        if (G.slots == 2) keys[at + 1] = (unsigned)k1;
    } else {
        G.keys[at] = k0;
        if (G.slots == 2) G.keys[at + 1] = k1;
    }
    G.vals[at] = 2 * p;
    if (G.slots == 2) G.vals[at + 1] = 2 * p + 1;
}

// Sorted key j (stored as uint32 when the keys fit 32 bits).
__device__ inline long long sorted_key(const Group& G, long long j) {
    return G.key32 ? (long long)reinterpret_cast<const unsigned*>(G.sorted_keys)[j]
                   : G.sorted_keys[j];
}

__device__ inline bool is_sentinel(const Group& G, long long key) {
    return (key & G.key_mask) == G.key_mask;
}

__global__ void __launch_bounds__(kThreads) o1_group_flags(Group G, long long total) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > total) return;
    if (j == total) { G.group_flag[j] = 0; return; }
    long long seg = (long long)G.slots * G.n_pen;
    long long k = sorted_key(G, j);
    bool start = !is_sentinel(G, k) && (j % seg == 0 || sorted_key(G, j - 1) != k);
    G.group_flag[j] = start ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) o1_group_starts(Group G, long long total, int work) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > total) return;
    long long seg = (long long)G.slots * G.n_pen;
    if (j == total) {
        int groups = G.group_scan[j];
        G.info[0] = groups;
        G.om_groups[work] = groups;
        return;
    }
    if (j % seg == 0) G.om_groups[j / seg] = G.group_scan[j];
    if (G.group_flag[j]) G.group_start[G.group_scan[j]] = (int)j;
    long long k = sorted_key(G, j);
    if (!is_sentinel(G, k) && (j + 1 == total || (j + 1) % seg == 0 || sorted_key(G, j + 1) != k))
        G.group_end[G.group_scan[j] + G.group_flag[j] - 1] = (int)(j + 1);
}

// Acceptance monomials c1^i c3^j in the order k = 0..P, i = 0..k (j = k - i);
// the host builds coef, ipow and jpow in this order.
template <int P>
struct Monomials {
    static constexpr int N = (P + 1) * (P + 2) / 2;
};

template <int P>
__device__ inline void fill_moments(double weight0, double weight2, double c1, double c3,
                                    int averaged, double* out0, double* out2) {
    double cp[P + 1], sp[P + 1];
    cp[0] = 1.0; sp[0] = 1.0;
    double perp2 = fmax(1.0 - c1 * c1, 0.0);
#pragma unroll
    for (int p = 1; p <= P; ++p) {
        cp[p] = cp[p - 1] * c1;
        if (averaged) sp[p] = (p % 2 == 1) ? 0.0 : sp[p - 2 >= 0 ? p - 2 : 0] * perp2 * (p - 1.0) / p;
        else sp[p] = sp[p - 1] * c3;
    }
    int m = 0;
#pragma unroll
    for (int k = 0; k <= P; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i, ++m) {
            out0[m] = weight0 * cp[i] * sp[k - i];
            out2[m] = weight2 * cp[i] * sp[k - i];
        }
    }
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}


constexpr int kPieceEntries = 64;    // group members per wave of the group sums

// Pieces of every group: ceil(members / kPieceEntries).
// Members of every group (groups + 1 entries, the last 0), scanned into
// member_off: the groups' members numbered consecutively.
__global__ void __launch_bounds__(kThreads) o1_member_counts(Group G, int groups, int* counts) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > groups) return;
    counts[g] = g == groups ? 0 : G.group_end[g] - G.group_start[g];
}

// Pieces of every group: the tiles of kPieceEntries consecutive members it touches.
__global__ void __launch_bounds__(kThreads) o1_piece_counts(Group G, int groups, int* counts) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > groups) return;
    if (g == groups) { counts[g] = 0; return; }
    int first = G.member_off[g], last = G.member_off[g + 1] - 1;
    counts[g] = last / kPieceEntries - first / kPieceEntries + 1;
}

// The group of every tile's first member: tile t (members 64t ..) starts in
// group g when member_off[g] <= 64t < member_off[g + 1]. One thread per group
// writes its tiles, so the group sums need no search.
__global__ void __launch_bounds__(kThreads) o1_tile_groups(Group G, int groups) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    long long t0 = ((long long)G.member_off[g] + kPieceEntries - 1) / kPieceEntries;
    long long t1 = ((long long)G.member_off[g + 1] + kPieceEntries - 1) / kPieceEntries;
    for (long long t = t0; t < t1; ++t) G.tile_group[t] = g;
}

// Group sums: one wave per tile of kPieceEntries consecutive members (of any
// number of groups), a lane per member; a fixed-order segmented scan across
// the lanes sums each group's members in the tile, and the group's last lane
// writes that piece. The pieces of a group are then added in piece order
// (deterministic). One tile per wave, no loop: a loop here makes the compiler
// keep about twice the registers live.
template <int P>
__global__ void __launch_bounds__(kThreads, 4) o1_group_sums(Group G, int groups) {
    constexpr int n_mono = Monomials<P>::N;
    constexpr int nv = 2 * n_mono + 6;
    // member rows, value-major with a padded stride: the lanes summing one
    // piece's values read one member at a time from distinct banks
    constexpr int kStride = kPieceEntries + 1;
    __shared__ int head_all[kThreads / 64][kPieceEntries];
    __shared__ int piece_lane_all[kThreads / 64][kPieceEntries + 1];
    __shared__ double rows_all[kThreads / 64][nv * kStride];
    int lane = threadIdx.x & 63;
    int* head = head_all[threadIdx.x >> 6];
    long long seg = (long long)G.slots * G.n_pen;
    long long members = G.member_off[groups];
    if (G.work_counts != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(G.work_counts + 3, (unsigned long long)members);
    long long tiles = (members + kPieceEntries - 1) / kPieceEntries;
    {
        // workgroups go to the 8 XCDs in turn; each XCD takes a contiguous
        // eighth of the tiles, so neighbouring OMs (which share pencils)
        // gather through one L2
        const long long used = (tiles + kThreads / 64 - 1) / (kThreads / 64);
        if ((long long)blockIdx.x >= used) return;
        const long long b = xcd_block(blockIdx.x, used);
        long long t = b * (kThreads / 64) + (threadIdx.x >> 6);
        if (t >= tiles) return;
        long long c0 = t * kPieceEntries;
        const int low = G.tile_group[t];   // member_off[low] <= c0 < member_off[low + 1]
        // the lane where each group of the tile starts holds the group's index;
        // a running maximum carries it to the group's other lanes
        wave_sync();
        head[lane] = -1;
        wave_sync();
        int gk = low + lane;
        if (gk < groups) {
            long long at = (long long)G.member_off[gk] - c0;
            if (lane == 0) at = 0;
            if (at < kPieceEntries) head[at] = gk;
        }
        wave_sync();
        int g = head[lane];
        for (int off = 1; off < 64; off <<= 1) {
            int other = __shfl_up(g, off, 64);
            if (lane >= off) g = max(g, other);
        }
        long long c = c0 + lane;
        bool valid = c < members;
        if (!valid) g = -1 - lane;      // a segment of its own
        double row[nv];
#pragma unroll
        for (int m = 0; m < nv; ++m) row[m] = 0.0;
        if (valid) {
            long long j = G.group_start[g] + (c - G.member_off[g]);
            int w = (int)(j / seg);
            int d = G.om.work[w];
            int v = G.sorted_vals[j];
            int p = v >> 1, node = v & 1;
            Entry x;
            pencil_entries(G, p, d, w, node, x, nullptr);
            double mom0[n_mono], mom2[n_mono];
            fill_moments<P>(x.w0, x.w2, x.c1, x.c3, x.averaged, mom0, mom2);
#pragma unroll
            for (int m = 0; m < n_mono; ++m) {
                row[m] = mom0[m];
                row[n_mono + m] = mom2[m];
            }
            double omega = fabs(mom0[0]);
            row[2 * n_mono] = omega;
            row[2 * n_mono + 1] = omega * x.theta;
            row[2 * n_mono + 2] = omega * x.r;
            row[2 * n_mono + 3] = omega * x.t_emit;
            row[2 * n_mono + 4] = omega * x.tau;
            row[2 * n_mono + 5] = omega * x.tau * x.tau;
        }
        // the tile's pieces (maximal runs of one group among the valid lanes,
        // which are a prefix of the wave) are summed from LDS: one lane per
        // (piece, value) adds the piece's members in lane order
        double* rows = rows_all[threadIdx.x >> 6];
#pragma unroll
        for (int m = 0; m < nv; ++m) rows[m * kStride + lane] = row[m];
        int* piece_lane = piece_lane_all[threadIdx.x >> 6];
        int* piece_group = head;   // the heads have been read
        const int prev = __shfl_up(g, 1, 64);
        const bool start = valid && (lane == 0 || prev != g);
        const unsigned long long starts = __ballot(start);
        const int n_pieces = __popcll(starts);
        const int n_valid = __popcll(__ballot(valid));
        if (start) {
            const int q = __popcll(starts & ((1ULL << lane) - 1ULL));
            piece_lane[q] = lane;
            piece_group[q] = g;
        }
        if (lane == 0) piece_lane[n_pieces] = n_valid;
        wave_sync();
        for (int task = lane; task < n_pieces * nv; task += 64) {
            const int q = task / nv, m = task - q * nv;
            const int a = piece_lane[q], b = piece_lane[q + 1];
            const int gq = piece_group[q];
            // four running sums (the loads issue back to back), combined in a
            // fixed order
            const double* r = rows + m * kStride;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int i = a;
            for (; i + 3 < b; i += 4) {
                s0 += r[i];
                s1 += r[i + 1];
                s2 += r[i + 2];
                s3 += r[i + 3];
            }
            for (; i < b; ++i) s0 += r[i];
            const double sum = (s0 + s1) + (s2 + s3);
            // a group within one tile is complete: its row goes straight to values
            const int first = G.piece_start[gq];
            double* out = G.piece_start[gq + 1] - first == 1
                              ? G.values + (size_t)gq * nv
                              : G.piece_values
                                    + (size_t)(first + (t - G.member_off[gq] / kPieceEntries)) * nv;
            out[m] = sum;
        }
    }
}

// Pieces of a group added in piece order, one thread per (tile, value): a
// group of more than one piece is combined by the tile after its first one
// (the tile whose first member lies in the group one tile past the group's
// first member's tile), so the launch covers tiles, not groups.
__global__ void __launch_bounds__(kThreads) o1_group_combine(Group G, int groups) {
    const int nv = 2 * G.opt.n_mono + 6;
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long members = G.member_off[groups];
    const long long tiles = (members + kPieceEntries - 1) / kPieceEntries;
    if (e >= tiles * nv) return;
    long long t = e / nv;
    int m = (int)(e - t * nv);
    if (t == 0) return;
    int g = G.tile_group[t];
    if (G.member_off[g] / kPieceEntries + 1 != t) return;
    int first = G.piece_start[g], last = G.piece_start[g + 1];
    double total = G.piece_values[(size_t)first * nv + m];
    for (int q = first + 1; q < last; ++q) total += G.piece_values[(size_t)q * nv + m];
    G.values[(size_t)g * nv + m] = total;
}

// Batches: consecutive groups of one OM with one key / theta_bins.
__global__ void __launch_bounds__(kThreads) o1_batch_flags(Group G, int groups, int work) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > groups) return;
    if (g == groups) { G.batch_flag[g] = 0; return; }
    long long seg = (long long)G.slots * G.n_pen;
    int w = (int)(G.group_start[g] / seg);
    long long key = (sorted_key(G, G.group_start[g]) & G.key_mask) / G.theta_bins;
    bool first = g == G.om_groups[w];
    G.batch_flag[g] = (first || (sorted_key(G, G.group_start[g - 1]) & G.key_mask)
                                        / G.theta_bins != key) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) o1_batch_starts(Group G, int groups, int work) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= groups && (g == groups || G.batch_flag[g])) G.batch_start[G.batch_scan[g]] = g;
    if (g <= work) G.om_batches[g] = G.batch_scan[G.om_groups[g]];
}

// ---- batch core

// The first N terms z^n / (n! (n + j + 1)) of i_j, j = 0, 1, 2 (Horner).
template <int N>
__device__ __forceinline__ void taylor_terms(double z, double* i0, double* i1, double* i2) {
    double a = kTaylor[0][N - 1], b = kTaylor[1][N - 1], c = kTaylor[2][N - 1];
#pragma unroll
    for (int n = N - 2; n >= 0; --n) {
        a = a * z + kTaylor[0][n];
        b = b * z + kTaylor[1][n];
        c = c * z + kTaylor[2][n];
    }
    *i0 = a; *i1 = b; *i2 = c;
}

// Inside |z| < 1 the series; the terms beyond the first N are below the FP64
// rounding of the sum (>= 1/3) when |z| < 1/16 and N = 9 (remainder < 1e-17)
// or |z| < 1/4 and N = 12 (< 1e-17), else all 24 are taken. The wave takes
// the longest series any of its lanes needs, so the coefficients stay
// uniform loads.
__device__ void ifuncs(double z, double ez, double* i0, double* i1, double* i2) {
    if (fabs(z) < 1.0) {
        if (__ballot(fabs(z) >= 0.25) != 0) taylor_terms<24>(z, i0, i1, i2);
        else if (__ballot(fabs(z) >= 0.0625) != 0) taylor_terms<12>(z, i0, i1, i2);
        else taylor_terms<9>(z, i0, i1, i2);
        return;
    }
    *i0 = (ez - 1.0) / z;
    *i1 = (ez * (z - 1.0) + 1.0) / (z * z);
    *i2 = (ez * (z * z - 2.0 * z + 2.0) - 2.0) / (z * z * z);
}

__device__ double quad_integral(double e_start, double alpha, double beta, double gamma,
                                double width, double kappa, double e_end) {
    double z = kappa * width;
    double ez = e_start > 0.0 ? e_end / e_start : exp(z);
    double i0, i1, i2;
    ifuncs(z, ez, &i0, &i1, &i2);
    return e_start * width * (alpha * i0 + width * (beta * i1 + width * gamma * i2));
}

__device__ int find_bin(const double* edges, int nb, double t, double inverse_width = 0.0) {
    if (t < edges[0]) return -1;
    if (t >= edges[nb]) return nb;
    if (inverse_width > 0.0) {
        // equal bins: start at the arithmetic guess and step to the edges' bin
        // (the same bin as the search below: edges[k] <= t < edges[k + 1])
        int k = (int)((t - edges[0]) * inverse_width);
        k = min(max(k, 0), nb - 1);
        for (int step = 0; step < 4; ++step) {
            if (t < edges[k]) --k;
            else if (t >= edges[k + 1]) ++k;
            else return k;
        }
    }
    int low = 0, high = nb;
    while (high - low > 1) {
        int middle = (low + high) / 2;
        if (t < edges[middle]) high = middle; else low = middle;
    }
    return low;
}

struct LambdaBins {
    double inverse_width;  // equal bins: 1 / width, else 0
    const double* edges;
    const double* eedge;   // (nb + 1) of this wavelength
    double* acc;           // (nb, 4)
    double* pre;           // (3)
    int nb, direct_bins;
};

__device__ void cubic_coefficients(const Group& G, const double* mom, double* a, double* b) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int m = 0; m < G.opt.n_mono; ++m) {
        double value = G.opt.coef[m] * mom[m];
        int i = G.opt.ipow[m], j = G.opt.jpow[m];
        if (j == 0) {
            if (i == 0) a0 += value; else if (i == 1) a1 += value;
            else if (i == 2) a2 += value; else a3 += value;
        } else if (j == 1) {
            if (i == 0) b0 += value; else if (i == 1) b1 += value; else b2 += value;
        } else if (j == 2) {
            if (i == 0) { a0 += value; a2 -= value; } else { a1 += value; a3 -= value; }
        } else {
            b0 += value; b2 -= value;
        }
    }
    a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3; b[0] = b0; b[1] = b1; b[2] = b2;
}

constexpr int kBatchGroups = 64;   // groups staged in LDS per profile pass
constexpr int kMaxDirect = 4;      // direct_bins limit of the panel events
static_assert(kMaxDirect >= 4, "prefix panels keep A, B, C and I_start in v[0..3]");
constexpr int kBatchWaves = 8;     // waves of a batch workgroup (one wavelength each)
constexpr int kBatchThreads = 64 * kBatchWaves;

// Deposits of one panel (_deposit_panel, _prompt_numba.py:186-253) as events:
// kind 0 none; kind 1 direct, v[e] added to bin k_lo + e for the set bits of
// mask; kind 2 prefix, (A, B, C, I_start) = v[0..3] added at k_lo (or A, B, C
// to the pre-window sums when pre is set) and (A, B, C, I_end = v_end)
// subtracted at k_end when end is set.
struct PanelEvents {
    int kind, k_lo, k_end, n, mask, pre, end;
    double v[kMaxDirect];   // direct: one integral per bin; prefix: A, B, C, I_start
    double v_end;           // prefix: I_end
};

__device__ __forceinline__ void panel_events(const LambdaBins& B, double t_a, double t_b, double e_a, double e_b,
                             double alpha, double beta, double gamma, double kappa,
                             PanelEvents& out) {
    out.kind = 0;
    int nb = B.nb;
    const double* edges = B.edges;
    double lo = t_a > edges[0] ? t_a : edges[0];
    double hi = t_b < edges[nb] ? t_b : edges[nb];
    if (hi <= lo) return;
    int k_lo = find_bin(edges, nb, lo, B.inverse_width);
    int k_hi = find_bin(edges, nb, hi, B.inverse_width);
    if (k_hi >= nb || (k_hi > k_lo && hi <= edges[k_hi])) k_hi -= 1;
    out.k_lo = k_lo;
    if (k_hi - k_lo + 1 <= B.direct_bins) {
        out.kind = 1;
        out.n = k_hi - k_lo + 1;
        out.mask = 0;
#pragma unroll
        for (int e = 0; e < kMaxDirect; ++e) {
            if (e >= out.n) break;
            int k = k_lo + e;
            double s_lo = lo > edges[k] ? lo : edges[k];
            double s_hi = hi < edges[k + 1] ? hi : edges[k + 1];
            if (s_hi <= s_lo) continue;
            double e_lo = s_lo == t_a ? e_a : B.eedge[k];
            if (s_lo != t_a && s_lo != edges[k]) e_lo = exp(kappa * s_lo);
            double e_hi = s_hi == t_b ? e_b : B.eedge[k + 1];
            if (s_hi != t_b && s_hi != edges[k + 1]) e_hi = exp(kappa * s_hi);
            double shift = s_lo - t_a;
            double a_loc = alpha + shift * (beta + shift * gamma);
            double b_loc = beta + 2.0 * gamma * shift;
            out.v[e] = quad_integral(e_lo, a_loc, b_loc, gamma, s_hi - s_lo, kappa, e_hi);
            out.mask |= 1 << e;
        }
        return;
    }
    out.kind = 2;
    double big_a = alpha - t_a * (beta - gamma * t_a);
    double big_b = beta - 2.0 * gamma * t_a;
    out.v[0] = big_a;
    out.v[1] = big_b;
    out.v[2] = gamma;
    out.pre = t_a < edges[0];
    if (!out.pre)
        out.v[3] = quad_integral(e_a, alpha, beta, gamma, edges[k_lo + 1] - t_a, kappa,
                                 B.eedge[k_lo + 1]);
    out.end = t_b < edges[nb];
    out.k_end = k_hi;
    if (out.end) {
        int k = find_bin(edges, nb, t_b, B.inverse_width);
        out.k_end = k;
        double shift = t_b - t_a;
        double a_loc = alpha + shift * (beta + shift * gamma);
        double b_loc = beta + 2.0 * gamma * shift;
        out.v_end = quad_integral(e_b, a_loc, b_loc, gamma, edges[k + 1] - t_b, kappa,
                                  B.eedge[k + 1]);
    }
}

// One round of deposits: every lane holds the events of one panel (panels in
// time order across the lanes). The events are flattened in (panel, bin) order,
// along which the bins never decrease, so every bin's events are contiguous and
// in panel order; one lane per bin adds them, the order of the reference's
// serial loop. Returns the number of bin events.
__device__ __forceinline__ int deposit_round(const PanelEvents& ev, int lane, const LambdaBins& B, int* ev_bin,
                             double (*ev_val)[4], int* head_at) {
    int mine = 0;
    if (ev.kind == 1) mine = __popc(ev.mask);
    else if (ev.kind == 2) mine = (ev.pre ? 0 : 1) + (ev.end ? 1 : 0);
    int offset = mine;
    for (int off = 1; off < 64; off <<= 1) {
        int other = __shfl_up(offset, off, 64);
        if (lane >= off) offset += other;
    }
    int n_events = __shfl(offset, 63, 64);
    offset -= mine;
    if (ev.kind == 1) {
#pragma unroll
        for (int e = 0; e < kMaxDirect; ++e) {
            if (!((ev.mask >> e) & 1)) continue;
            ev_bin[offset] = ev.k_lo + e;
            ev_val[offset][0] = 0.0; ev_val[offset][1] = 0.0;
            ev_val[offset][2] = 0.0; ev_val[offset][3] = ev.v[e];
            ++offset;
        }
    } else if (ev.kind == 2) {
        if (!ev.pre) {
            ev_bin[offset] = ev.k_lo;
            ev_val[offset][0] = ev.v[0]; ev_val[offset][1] = ev.v[1];
            ev_val[offset][2] = ev.v[2]; ev_val[offset][3] = ev.v[3];
            ++offset;
        }
        if (ev.end) {
            ev_bin[offset] = ev.k_end;
            ev_val[offset][0] = -ev.v[0]; ev_val[offset][1] = -ev.v[1];
            ev_val[offset][2] = -ev.v[2]; ev_val[offset][3] = -ev.v_end;
        }
    }
    wave_sync();
    // the first event of every bin, compacted in order
    int heads = 0;
    for (int c = 0; c < n_events; c += 64) {
        int i = c + lane;
        bool head = i < n_events && (i == 0 || ev_bin[i] != ev_bin[i - 1]);
        unsigned long long ballot = __ballot(head);
        if (head) head_at[heads + __popcll(ballot & ((1ULL << lane) - 1))] = i;
        heads += __popcll(ballot);
    }
    wave_sync();
    for (int h = lane; h < heads; h += 64) {
        int i0 = head_at[h], i1 = h + 1 < heads ? head_at[h + 1] : n_events;
        double* a = B.acc + 4 * (size_t)ev_bin[i0];
        double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        for (int i = i0; i < i1; ++i) {
            a0 += ev_val[i][0]; a1 += ev_val[i][1];
            a2 += ev_val[i][2]; a3 += ev_val[i][3];
        }
        a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3;
    }
    // window-front sums of the panels that start before the first edge, in
    // panel order (rare: a ballot finds them)
    unsigned long long pre = __ballot(ev.kind == 2 && ev.pre);
    for (unsigned long long rest = pre; rest != 0; rest &= rest - 1) {
        int q = __ffsll((long long)rest) - 1;
        double a = __shfl(ev.v[0], q, 64), b = __shfl(ev.v[1], q, 64),
               c = __shfl(ev.v[2], q, 64);
        if (lane == 0) { B.pre[0] += a; B.pre[1] += b; B.pre[2] += c; }
    }
    wave_sync();
    return n_events;
}

// ---- batches: header, path grid and profile, per-wavelength deposit

// A batch's moment means and path grid (_grouped_pencils_order1_block
// :1924-1960, _batch_grid :258-286). n = 0 marks a batch that carries no light.
struct BatchHead {
    double r_b, t_b, tau_var, sigma, u_win, u_end;
    int n, n_in, n_out, backward;
    int start, stop, pad0, pad1;
};

__device__ __forceinline__ BatchHead batch_head(const Group& G, int b, int lane) {
    BatchHead h;
    int start = G.batch_start[b], stop = G.batch_start[b + 1];
    int n_mono = G.opt.n_mono;
    int nv = 2 * n_mono + 6;
    int nb = G.opt.bins;
    h.start = start;
    h.stop = stop;
    h.pad0 = h.pad1 = 0;
    h.n = h.n_in = h.n_out = 0;
    h.r_b = h.t_b = h.tau_var = h.sigma = h.u_win = h.u_end = 0.0;
    h.backward = ((sorted_key(G, G.group_start[start]) & G.key_mask) / G.theta_bins) % 2 == 1;
    double omega_sum = 0, r_sum = 0, t_sum = 0, tau_sum = 0, tau2_sum = 0, theta_small = M_PI;
    for (int q = start + lane; q < stop; q += 64) {
        const double* row = G.values + (size_t)q * nv;
        double omega = row[2 * n_mono];
        if (omega > 0.0) {
            omega_sum += omega;
            r_sum += row[2 * n_mono + 2];
            t_sum += row[2 * n_mono + 3];
            tau_sum += row[2 * n_mono + 4];
            tau2_sum += row[2 * n_mono + 5];
            theta_small = fmin(theta_small, row[2 * n_mono + 1] / omega);
        }
    }
    omega_sum = wave_sum(omega_sum);
    r_sum = wave_sum(r_sum);
    t_sum = wave_sum(t_sum);
    tau_sum = wave_sum(tau_sum);
    tau2_sum = wave_sum(tau2_sum);
    theta_small = wave_min(theta_small);
    if (!(omega_sum > 0.0)) return h;
    double r_b = r_sum / omega_sum;
    double t_b = t_sum / omega_sum;
    double tau_mean = tau_sum / omega_sum;
    double tau_var = fmax(tau2_sum / omega_sum - tau_mean * tau_mean, 0.0);
    theta_small = fmax(theta_small, 1e-9);
    double sigma;
    if (h.backward)
        sigma = fmax(0.5 * r_b * sin(theta_small) * G.opt.y_g,
                     0.25 * r_b * (1.0 - cos(theta_small)));
    else
        sigma = fmax(0.5 * r_b * sin(theta_small)
                         * fmax(G.opt.y_g, 0.5 * tan(0.5 * theta_small)),
                     1e-9 * r_b);
    double h_in = G.opt.path_step;
    double h_out = G.screen ? G.opt.path_step : G.opt.path_step_tail;
    double ext_end = G.opt.tail_exponent / G.mut_min;
    double ext_win = G.screen ? ext_end : G.speed_max * (G.opt.edges[nb] - t_b) - r_b;
    if (ext_win < 0.0) ext_win = 0.0;
    if (ext_win > ext_end) ext_win = ext_end;
    double u_end = asinh(ext_end / sigma);
    double u_win = asinh(ext_win / sigma);
    int n_in = u_win > 0.0 ? (int)ceil(u_win / h_in) : 0;
    int n_out = u_end > u_win ? (int)ceil((u_end - u_win) / h_out) : 0;
    int capacity = min(G.opt.max_path_nodes, kMaxPathNodes);
    if (n_in + n_out + 1 > capacity) {
        double scale = (capacity - 1.0) / (n_in + n_out);
        n_in = (int)(n_in * scale);
        n_out = capacity - 1 - n_in;
    }
    h.r_b = r_b;
    h.t_b = t_b;
    h.tau_var = tau_var;
    h.sigma = sigma;
    h.u_win = u_win;
    h.u_end = u_end;
    h.n_in = n_in;
    h.n_out = n_out;
    h.n = n_in + n_out + 1;
    return h;
}

// Path node k: asinh-spaced in the window, then with the tail step.
__device__ inline double path_node(const BatchHead& h, int k) {
    if (k == 0) return h.r_b;
    if (k <= h.n_in) return h.r_b + h.sigma * sinh(h.u_win * k / h.n_in);
    return h.r_b + h.sigma * sinh(h.u_win + (h.u_end - h.u_win) * (k - h.n_in) / h.n_out);
}

// A group's direction and (for degree <= 3) its cubic acceptance coefficients:
// A0[4] B0[3] A2[4] B2[3] sine cosine; a negative sine marks an empty group.
template <int P>
__device__ inline void stage_group(const Group& G, const double* row, double* cf) {
    int n_mono = G.opt.n_mono;
    double omega = row[2 * n_mono];
    if (omega <= 0.0) { cf[14] = -1.0; return; }
    double theta = fmax(row[2 * n_mono + 1] / omega, 1e-9);
    cf[14] = sin(theta);
    cf[15] = cos(theta);
    if (P <= 3) {
        cubic_coefficients(G, row, cf, cf + 4);
        cubic_coefficients(G, row + n_mono, cf + 7, cf + 11);
    }
}

// One group's HG single-scattering term at path node s_k, added to the two
// profile sums (_batch_profile, :300-370).
template <int P>
__device__ inline void profile_add(const Group& G, const double* cf, const double* row,
                                   double s_k, double r_b, int backward, double gg, double g1,
                                   double norm, double& acc0, double& acc2) {
    double sin_t = cf[14];
    if (sin_t < 0.0) return;
    double cos_t = cf[15];
    double rs = r_b * sin_t;
    double c, sn, path_jacobian;
    if (backward) {
        double shift = s_k - r_b * cos_t;
        double denominator = rs * rs + shift * shift;
        c = (rs * rs - shift * shift) / denominator;
        sn = 2.0 * rs * shift / denominator;
        path_jacobian = 2.0 * r_b / denominator;
    } else {
        double y_theta = sin_t / (1.0 + cos_t);
        double jac = 2.0 / (r_b * sin_t);
        double y = (s_k - r_b) / rs + y_theta;
        double inv = 1.0 / (1.0 + y * y);
        c = (1.0 - y * y) * inv;
        sn = 2.0 * y * inv;
        path_jacobian = jac * inv;
    }
    double a0, a2;
    if (P <= 3) {
        a0 = cf[0] + c * (cf[1] + c * (cf[2] + c * cf[3]));
        a0 += sn * (cf[4] + c * (cf[5] + c * cf[6]));
        a2 = cf[7] + c * (cf[8] + c * (cf[9] + c * cf[10]));
        a2 += sn * (cf[11] + c * (cf[12] + c * cf[13]));
    } else {
        double cpow[P + 1], spow[P + 1];
        cpow[0] = 1.0; spow[0] = 1.0;
#pragma unroll
        for (int pw = 1; pw <= P; ++pw) {
            cpow[pw] = cpow[pw - 1] * c;
            spow[pw] = spow[pw - 1] * sn;
        }
        a0 = 0.0; a2 = 0.0;
        int m = 0;
#pragma unroll
        for (int kk = 0; kk <= P; ++kk) {
#pragma unroll
            for (int i = 0; i <= kk; ++i, ++m) {
                double basis = G.opt.coef[m] * cpow[i] * spow[kk - i];
                a0 += basis * row[m];
                a2 += basis * row[Monomials<P>::N + m];
            }
        }
    }
    double den = g1 - gg * c;
    double value = path_jacobian * norm / (den * sqrt(den));
    acc0 += a0 * value;
    acc2 += a2 * value;
}

// _batch_deposit (:451-551) of one batch at one wavelength: the panels are
// spread over the lanes 64 at a time and every bin adds its panels' events in
// panel order (the order of the reference's serial loop). acc and pre are the
// wavelength's accumulators. Returns the wavelength's charge of the batch (the
// same value in every lane).
__device__ __forceinline__ double deposit_lambda(const Group& G, int lam, const BatchHead& h, const double* path,
                                 const double* g0, const double* g2, double* acc, double* pre,
                                 int lane, int* ev_bin, double (*ev_val)[4], int* head_at,
                                 unsigned long long& count_panels,
                                 unsigned long long& count_events) {
    int nb = G.opt.bins;
    int n = h.n;
    double r_b = h.r_b, t_b = h.t_b;
    double v = G.opt.speed[lam];
    const double inv_v = 1.0 / v;   // one division per batch, products per panel
    double mut = G.opt.mut[lam];
    double kappa = -mut * v;
    // last = the serial walk-back `while (last > 2 && path[last - 1] >
    // limit) --last` from n - 1; path increases, so it is the count of
    // path[0 .. n - 2] within the limit, clamped to [2, n - 1]
    double limit = r_b + G.opt.tail_exponent / mut;
    int within = 0;
    for (int k0 = 0; k0 < n - 1; k0 += 64) {
        int k = k0 + lane;
        within += __popcll(__ballot(k < n - 1 && path[k] <= limit));
    }
    int last = min(max(within, 2), n - 1);
    if (last < 2) return 0.0;
    double spread = fmin(sqrt(12.0 * h.tau_var), (path[last] - r_b) * inv_v);
    double front_time = t_b + r_b * inv_v;
    double end_time = t_b + path[last] * inv_v;
    int outside = front_time - 0.5 * spread >= G.opt.edges[nb]
                  || end_time + 0.5 * spread <= G.opt.edges[0];
    double t_base = outside ? 0.0 : t_b;
    double e_emit = exp(kappa * t_base);
    double h_scale = G.opt.mus[lam] * v / e_emit;
    double w0 = G.opt.s0w[lam] * h_scale, w2 = G.opt.s2w[lam] * h_scale;
    LambdaBins B;
    B.inverse_width = G.opt.bin_inverse_width;
    B.edges = G.opt.edges;
    B.eedge = G.opt.eedge + (size_t)lam * (nb + 1);
    B.acc = acc;
    B.pre = pre;
    B.nb = nb;
    B.direct_bins = G.opt.direct_bins;
    bool bins = G.compute_bins && !outside;
    double h_front = w0 * g0[0] + w2 * g2[0];
    double e_first = e_emit * exp(-mut * path[0]);
    double e_last = e_emit * exp(-mut * path[last]);
    // panels: (2i, 2i+1, 2i+2) for i < last / 2; an odd last closes with
    // the one-step panel at j = last - 1
    int panels = last / 2 + (last & 1);
    double total_lane = 0.0;
    unsigned long long events_total = 0;
    double carry = e_first;   // e_b of the panel before the round's first lane
    for (int base = 0; base < panels; base += 64) {
        int i = base + lane;
        PanelEvents ev;
        ev.kind = 0;
        double t_a = 0.0, t_e = 0.0, alpha = 0.0, beta = 0.0, gamma = 0.0, e_b = 0.0;
        if (i < panels) {
            int j;
            if (i < last / 2) {
                j = 2 * i;
                t_a = t_base + path[j] * inv_v;
                double t_m = t_base + path[j + 1] * inv_v;
                t_e = t_base + path[j + 2] * inv_v;
                double h0 = (w0 * g0[j] + w2 * g2[j]) - h_front;
                double h1 = (w0 * g0[j + 1] + w2 * g2[j + 1]) - h_front;
                double h2 = (w0 * g0[j + 2] + w2 * g2[j + 2]) - h_front;
                double u_m = t_m - t_a, u_b = t_e - t_a;
                double d1 = (h1 - h0) / u_m;
                double d2 = ((h2 - h1) / (t_e - t_m) - d1) / u_b;
                alpha = h0;
                beta = d1 - d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-mut * path[j + 2]);
            } else {
                j = last - 1;
                double t_p = t_base + path[j - 1] * inv_v;
                t_a = t_base + path[j] * inv_v;
                t_e = t_base + path[j + 1] * inv_v;
                double hp = (w0 * g0[j - 1] + w2 * g2[j - 1]) - h_front;
                double h0 = (w0 * g0[j] + w2 * g2[j]) - h_front;
                double h1 = (w0 * g0[j + 1] + w2 * g2[j + 1]) - h_front;
                double u_m = t_a - t_p, u_b = t_e - t_p;
                double d1 = (h0 - hp) / u_m;
                double d2 = ((h1 - h0) / (t_e - t_a) - d1) / u_b;
                alpha = h0;
                beta = d1 - d2 * u_m + 2.0 * d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-mut * path[j + 1]);
            }
        }
        // a panel starts at the node its predecessor ends at, so its e_a is the
        // predecessor's e_b (the serial loop carries it the same way): from
        // the lane before, or from the previous round's last lane
        double before = __shfl_up(e_b, 1, 64);
        double e_a = lane == 0 ? carry : before;
        carry = __shfl(e_b, 63, 64);
        if (i < panels) {
            total_lane += quad_integral(e_a, alpha, beta, gamma, t_e - t_a, kappa, e_b);
            if (bins) panel_events(B, t_a, t_e, e_a, e_b, alpha, beta, gamma, kappa, ev);
        }
        if (!bins) continue;
        events_total += deposit_round(ev, lane, B, ev_bin, ev_val, head_at);
    }
    double total = wave_sum(total_lane);
    double t_f = t_base + path[0] * inv_v, t_end = t_base + path[last] * inv_v;
    double exact = quad_integral(e_first, h_front, 0.0, 0.0, t_end - t_f, kappa, e_last);
    total += exact;
    // the front ramp and flat tail after every panel, as one more round
    // (lanes 0 and 1; their times increase, so the round's bins do too)
    if (bins) {
        PanelEvents ev;
        ev.kind = 0;
        if (spread > 0.0) {
            double t_1 = t_f - 0.5 * spread, t_2 = t_f + 0.5 * spread;
            double e_1 = exp(kappa * t_1), e_2 = exp(kappa * t_2);
            double ramp = quad_integral(e_1, 0.0, 1.0 / spread, 0.0, spread, kappa, e_2);
            double flat = t_end > t_2
                              ? quad_integral(e_2, 1.0, 0.0, 0.0, t_end - t_2, kappa, e_last)
                              : 0.0;
            double height = ramp + flat != 0.0 ? exact / (ramp + flat) : 0.0;
            if (lane == 0)
                panel_events(B, t_1, t_2, e_1, e_2, 0.0, height / spread, 0.0, kappa, ev);
            else if (lane == 1 && t_end > t_2)
                panel_events(B, t_2, t_end, e_2, e_last, height, 0.0, 0.0, kappa, ev);
        } else if (lane == 0 && t_end > t_f) {
            panel_events(B, t_f, t_end, e_first, e_last, h_front, 0.0, 0.0, kappa, ev);
        }
        events_total += deposit_round(ev, lane, B, ev_bin, ev_val, head_at);
    }
    count_panels += (unsigned long long)panels;
    count_events += events_total;
    return total;
}

// deposit_round for two wavelengths at once, one per half of the wave (lanes
// 0-31 the low one, 32-63 the high one). Every half's events are in (panel,
// bin) order; the flattened events carry their half in the key, so every
// (wavelength, bin) is a contiguous run summed by one lane in panel order, as
// in deposit_round. Returns the number of bin events of both halves.
__device__ __forceinline__ int deposit_round_pair(const PanelEvents& ev, int lane,
                                                  double* acc_lo, double* acc_hi,
                                                  double* pre_lo, double* pre_hi, int* ev_key,
                                                  double (*ev_val)[4], int* head_at) {
    constexpr int kHalfBit = 1 << 30;
    const int half = lane >> 5;
    int mine = 0;
    if (ev.kind == 1) mine = __popc(ev.mask);
    else if (ev.kind == 2) mine = (ev.pre ? 0 : 1) + (ev.end ? 1 : 0);
    int offset = mine;
    for (int off = 1; off < 64; off <<= 1) {
        int other = __shfl_up(offset, off, 64);
        if (lane >= off) offset += other;
    }
    int n_events = __shfl(offset, 63, 64);
    offset -= mine;
    const int tag = half ? kHalfBit : 0;
    if (ev.kind == 1) {
#pragma unroll
        for (int e = 0; e < kMaxDirect; ++e) {
            if (!((ev.mask >> e) & 1)) continue;
            ev_key[offset] = tag | (ev.k_lo + e);
            ev_val[offset][0] = 0.0; ev_val[offset][1] = 0.0;
            ev_val[offset][2] = 0.0; ev_val[offset][3] = ev.v[e];
            ++offset;
        }
    } else if (ev.kind == 2) {
        if (!ev.pre) {
            ev_key[offset] = tag | ev.k_lo;
            ev_val[offset][0] = ev.v[0]; ev_val[offset][1] = ev.v[1];
            ev_val[offset][2] = ev.v[2]; ev_val[offset][3] = ev.v[3];
            ++offset;
        }
        if (ev.end) {
            ev_key[offset] = tag | ev.k_end;
            ev_val[offset][0] = -ev.v[0]; ev_val[offset][1] = -ev.v[1];
            ev_val[offset][2] = -ev.v[2]; ev_val[offset][3] = -ev.v_end;
        }
    }
    wave_sync();
    int heads = 0;
    for (int c = 0; c < n_events; c += 64) {
        int i = c + lane;
        bool head = i < n_events && (i == 0 || ev_key[i] != ev_key[i - 1]);
        unsigned long long ballot = __ballot(head);
        if (head) head_at[heads + __popcll(ballot & ((1ULL << lane) - 1))] = i;
        heads += __popcll(ballot);
    }
    wave_sync();
    for (int h = lane; h < heads; h += 64) {
        int i0 = head_at[h], i1 = h + 1 < heads ? head_at[h + 1] : n_events;
        const int key = ev_key[i0];
        double* a = ((key & kHalfBit) ? acc_hi : acc_lo) + 4 * (size_t)(key & (kHalfBit - 1));
        double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        for (int i = i0; i < i1; ++i) {
            a0 += ev_val[i][0]; a1 += ev_val[i][1];
            a2 += ev_val[i][2]; a3 += ev_val[i][3];
        }
        a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3;
    }
    // window-front sums of the panels that start before the first edge, each
    // half in panel order
    unsigned long long pre = __ballot(ev.kind == 2 && ev.pre);
    for (unsigned long long rest = pre; rest != 0; rest &= rest - 1) {
        int q = __ffsll((long long)rest) - 1;
        double a = __shfl(ev.v[0], q, 64), b = __shfl(ev.v[1], q, 64),
               c = __shfl(ev.v[2], q, 64);
        if (lane == 0) {
            double* p = q < 32 ? pre_lo : pre_hi;
            p[0] += a; p[1] += b; p[2] += c;
        }
    }
    wave_sync();
    return n_events;
}

__device__ __forceinline__ double half_sum(double v) {
    for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// deposit_lambda for two wavelengths of one batch at once, one per half of
// the wave (lam_hi < 0 or mus = 0 leaves a half idle): each half spreads its
// panels over its 32 lanes, 32 at a time, and both halves share every
// deposit round. Returns, in every lane, the charge of its half's wavelength.
__device__ __forceinline__ double deposit_pair(const Group& G, int lam_lo, int lam_hi,
                                               const BatchHead& h, const double* path,
                                               const double* g0, const double* g2,
                                               double* acc_lo, double* acc_hi, double* pre_lo,
                                               double* pre_hi, int lane, int* ev_key,
                                               double (*ev_val)[4], int* head_at,
                                               unsigned long long& count_panels,
                                               unsigned long long& count_events) {
    const int half = lane >> 5, sub = lane & 31;
    const int lam = half ? lam_hi : lam_lo;
    bool active = lam >= 0 && G.opt.mus[lam] != 0.0;
    const int lam_s = active ? lam : lam_lo;   // a valid index for the idle half
    int nb = G.opt.bins;
    int n = h.n;
    double r_b = h.r_b, t_b = h.t_b;
    double v = G.opt.speed[lam_s];
    const double inv_v = 1.0 / v;   // one division per batch, products per panel
    double mut = G.opt.mut[lam_s];
    double kappa = -mut * v;
    // last as in deposit_lambda, per half
    double limit = r_b + G.opt.tail_exponent / mut;
    int within = 0;
    for (int k0 = 0; k0 < n - 1; k0 += 32) {
        int k = k0 + sub;
        unsigned long long ballot = __ballot(active && k < n - 1 && path[k] <= limit);
        within += __popcll(half ? (ballot >> 32) : (ballot & 0xffffffffULL));
    }
    int last = min(max(within, 2), n - 1);
    if (last < 2) active = false;
    if (!active) last = 0;   // an idle half reads node 0 only (every batch has one)
    double spread = fmin(sqrt(12.0 * h.tau_var), (path[last] - r_b) * inv_v);
    double front_time = t_b + r_b * inv_v;
    double end_time = t_b + path[last] * inv_v;
    int outside = front_time - 0.5 * spread >= G.opt.edges[nb]
                  || end_time + 0.5 * spread <= G.opt.edges[0];
    double t_base = outside ? 0.0 : t_b;
    double e_emit = exp(kappa * t_base);
    double h_scale = G.opt.mus[lam_s] * v / e_emit;
    double w0 = G.opt.s0w[lam_s] * h_scale, w2 = G.opt.s2w[lam_s] * h_scale;
    LambdaBins B;
    B.inverse_width = G.opt.bin_inverse_width;
    B.edges = G.opt.edges;
    B.eedge = G.opt.eedge + (size_t)lam_s * (nb + 1);
    B.acc = half ? acc_hi : acc_lo;
    B.pre = half ? pre_hi : pre_lo;
    B.nb = nb;
    B.direct_bins = G.opt.direct_bins;
    bool bins = active && G.compute_bins && !outside;
    bool bins_any = __ballot(bins) != 0;
    double h_front = w0 * g0[0] + w2 * g2[0];
    double e_first = e_emit * exp(-mut * path[0]);
    double e_last = e_emit * exp(-mut * path[last]);
    int panels = active ? last / 2 + (last & 1) : 0;
    int panels_max = max(__shfl(panels, 0, 64), __shfl(panels, 32, 64));
    double total_lane = 0.0;
    unsigned long long events_total = 0;
    double carry = e_first;   // e_b of the panel before the half's round's first lane
    for (int base = 0; base < panels_max; base += 32) {
        int i = base + sub;
        PanelEvents ev;
        ev.kind = 0;
        double t_a = 0.0, t_e = 0.0, alpha = 0.0, beta = 0.0, gamma = 0.0, e_b = 0.0;
        if (i < panels) {
            int j;
            if (i < last / 2) {
                j = 2 * i;
                t_a = t_base + path[j] * inv_v;
                double t_m = t_base + path[j + 1] * inv_v;
                t_e = t_base + path[j + 2] * inv_v;
                double h0 = (w0 * g0[j] + w2 * g2[j]) - h_front;
                double h1 = (w0 * g0[j + 1] + w2 * g2[j + 1]) - h_front;
                double h2 = (w0 * g0[j + 2] + w2 * g2[j + 2]) - h_front;
                double u_m = t_m - t_a, u_b = t_e - t_a;
                double d1 = (h1 - h0) / u_m;
                double d2 = ((h2 - h1) / (t_e - t_m) - d1) / u_b;
                alpha = h0;
                beta = d1 - d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-mut * path[j + 2]);
            } else {
                j = last - 1;
                double t_p = t_base + path[j - 1] * inv_v;
                t_a = t_base + path[j] * inv_v;
                t_e = t_base + path[j + 1] * inv_v;
                double hp = (w0 * g0[j - 1] + w2 * g2[j - 1]) - h_front;
                double h0 = (w0 * g0[j] + w2 * g2[j]) - h_front;
                double h1 = (w0 * g0[j + 1] + w2 * g2[j + 1]) - h_front;
                double u_m = t_a - t_p, u_b = t_e - t_p;
                double d1 = (h0 - hp) / u_m;
                double d2 = ((h1 - h0) / (t_e - t_a) - d1) / u_b;
                alpha = h0;
                beta = d1 - d2 * u_m + 2.0 * d2 * u_m;
                gamma = d2;
                e_b = e_emit * exp(-mut * path[j + 1]);
            }
        }
        // e_a from the predecessor panel, as in deposit_lambda (per half)
        double before = __shfl_up(e_b, 1, 64);
        double e_a = sub == 0 ? carry : before;
        carry = __shfl(e_b, (lane & 32) | 31, 64);
        if (i < panels) {
            total_lane += quad_integral(e_a, alpha, beta, gamma, t_e - t_a, kappa, e_b);
            if (bins) panel_events(B, t_a, t_e, e_a, e_b, alpha, beta, gamma, kappa, ev);
        }
        if (!bins_any) continue;
        events_total += deposit_round_pair(ev, lane, acc_lo, acc_hi, pre_lo, pre_hi, ev_key,
                                           ev_val, head_at);
    }
    double total = half_sum(total_lane);
    double t_f = t_base + path[0] * inv_v, t_end = t_base + path[last] * inv_v;
    double exact = quad_integral(e_first, h_front, 0.0, 0.0, t_end - t_f, kappa, e_last);
    total += exact;
    // the front ramp and flat tail after every panel, as one more round
    // (lanes 0 and 1 of each half)
    if (bins_any) {
        PanelEvents ev;
        ev.kind = 0;
        if (bins) {
            if (spread > 0.0) {
                double t_1 = t_f - 0.5 * spread, t_2 = t_f + 0.5 * spread;
                double e_1 = exp(kappa * t_1), e_2 = exp(kappa * t_2);
                double ramp = quad_integral(e_1, 0.0, 1.0 / spread, 0.0, spread, kappa, e_2);
                double flat = t_end > t_2
                                  ? quad_integral(e_2, 1.0, 0.0, 0.0, t_end - t_2, kappa,
                                                  e_last)
                                  : 0.0;
                double height = ramp + flat != 0.0 ? exact / (ramp + flat) : 0.0;
                if (sub == 0)
                    panel_events(B, t_1, t_2, e_1, e_2, 0.0, height / spread, 0.0, kappa, ev);
                else if (sub == 1 && t_end > t_2)
                    panel_events(B, t_2, t_end, e_2, e_last, height, 0.0, 0.0, kappa, ev);
            } else if (sub == 0 && t_end > t_f) {
                panel_events(B, t_f, t_end, e_first, e_last, h_front, 0.0, 0.0, kappa, ev);
            }
        }
        events_total += deposit_round_pair(ev, lane, acc_lo, acc_hi, pre_lo, pre_hi, ev_key,
                                           ev_val, head_at);
    }
    count_panels += (unsigned long long)(__shfl(panels, 0, 64) + __shfl(panels, 32, 64));
    count_events += events_total;
    return active ? total : 0.0;
}

// The batch kernels run in one of two forms. The wide form (when the caller's
// scratch holds every batch's path nodes): o1_batch_heads (one wave per
// batch: moment means and node count), a scan of the node counts into the
// node pool, o1_batch_profile (one wave per batch: path nodes and profile
// sums into the pool) and o1_batch_deposit (one wave per (OM, segment,
// wavelength)). Otherwise o1_batches does all of it with one workgroup per
// (OM, segment) and the path in LDS. Both sum every slot's batches in the same
// order.
constexpr int kWideWaves = 4;
constexpr int kWideThreads = 64 * kWideWaves;
constexpr int kWideMaxBlocks = 2048;   // grid-stride bound of the per-batch kernels

__device__ inline bool wide_form(const Group& G, int groups) {
    return G.node_off[groups] <= G.pool_nodes;
}

// Headers and node counts of every batch; node_n is 0 from the last batch to groups.
__global__ void __launch_bounds__(kWideThreads) o1_batch_heads(Group G, int groups) {
    int lane = threadIdx.x & 63;
    int wave = blockIdx.x * kWideWaves + (threadIdx.x >> 6);
    int waves = gridDim.x * kWideWaves;
    int batches = G.om_batches[G.om.work_count];
    for (int b = wave; b <= groups; b += waves) {
        if (b < batches) {
            BatchHead h = batch_head(G, b, lane);
            if (lane == 0) {
                G.heads[b] = h;
                G.node_n[b] = h.n;
            }
        } else if (lane == 0) {
            G.node_n[b] = 0;
        }
    }
}

// Path nodes and the profile of every batch, one wave per batch: each lane
// owns nodes lane, lane + 64, ... and sums the batch's groups in key order.
template <int P>
__global__ void __launch_bounds__(kWideThreads) o1_batch_profile(Group G, int groups) {
    __shared__ double coeff_all[kWideWaves][kBatchGroups][16];
    __shared__ unsigned long long s_terms;
    int tid = threadIdx.x, lane = tid & 63;
    int wave = blockIdx.x * kWideWaves + (tid >> 6);
    int waves = gridDim.x * kWideWaves;
    double (*coeff)[16] = coeff_all[tid >> 6];
    if (tid == 0) s_terms = 0;
    __syncthreads();
    bool wide = wide_form(G, groups);
    int batches = G.om_batches[G.om.work_count];
    int nv = 2 * G.opt.n_mono + 6;
    double gg = 2.0 * G.opt.g, g1 = 1.0 + G.opt.g * G.opt.g;
    double norm = (1.0 - G.opt.g * G.opt.g) / (4.0 * M_PI);
    unsigned long long terms = 0;
    for (int b = wave; wide && b < batches; b += waves) {
        BatchHead h = G.heads[b];
        if (h.n == 0) continue;
        int n = h.n;
        size_t off = (size_t)G.node_off[b];
        double* path = G.pool_path + off;
        double* g0 = G.pool_g0 + off;
        double* g2 = G.pool_g2 + off;
        for (int k = lane; k < n; k += 64) path[k] = path_node(h, k);
        terms += (unsigned long long)n * (unsigned long long)(h.stop - h.start);
        for (int c0 = h.start; c0 < h.stop; c0 += kBatchGroups) {
            int chunk = min(kBatchGroups, h.stop - c0);
            wave_sync();   // the previous chunk's coefficients have been read
            for (int q = lane; q < chunk; q += 64)
                stage_group<P>(G, G.values + (size_t)(c0 + q) * nv, coeff[q]);
            wave_sync();
            for (int k = lane; k < n; k += 64) {
                double acc0 = c0 == h.start ? 0.0 : g0[k];
                double acc2 = c0 == h.start ? 0.0 : g2[k];
                double s_k = path_node(h, k);
                for (int q = 0; q < chunk; ++q)
                    profile_add<P>(G, coeff[q], G.values + (size_t)(c0 + q) * nv, s_k, h.r_b,
                                   h.backward, gg, g1, norm, acc0, acc2);
                g0[k] = acc0;
                g2[k] = acc2;
            }
        }
    }
    if (G.work_counts != nullptr) {
        if (lane == 0 && terms) atomicAdd(&s_terms, terms);
        __syncthreads();
        if (tid == 0 && s_terms) atomicAdd(G.work_counts, s_terms);
    }
}

// One wave per (OM, segment, wavelength): the segment's batches b = first + s,
// first + s + S, ... deposited in order into the slot's accumulators of that
// wavelength.
__global__ void __launch_bounds__(kWideThreads) o1_batch_deposit(Group G, int groups) {
    __shared__ int ev_bin_all[kWideWaves][64 * kMaxDirect];
    __shared__ double ev_val_all[kWideWaves][64 * kMaxDirect][4];
    __shared__ int head_at_all[kWideWaves][64 * kMaxDirect];
    __shared__ unsigned long long s_counts[2];
    int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid < 2) s_counts[tid] = 0;
    __syncthreads();
    int nl = G.opt.bands, nb = G.opt.bins, S = G.segments;
    int pairs = (nl + 1) / 2;
    long long gw = (long long)blockIdx.x * kWideWaves + wv;
    long long slots = (long long)G.om.work_count * S;
    unsigned long long count_panels = 0, count_events = 0;
    if (wide_form(G, groups) && gw < slots * pairs) {
        int pair = (int)(gw % pairs);
        long long slot = gw / pairs;
        int w = (int)(slot / S), seg = (int)(slot % S);
        int lam_lo = 2 * pair, lam_hi = 2 * pair + 1 < nl ? 2 * pair + 1 : -1;
        double* acc_lo = G.part_acc + ((size_t)slot * nl + lam_lo) * nb * 4;
        double* acc_hi = lam_hi >= 0 ? acc_lo + (size_t)nb * 4 : acc_lo;
        double* pre_lo = G.part_pre + (size_t)slot * nl * 3 + 3 * lam_lo;
        double* pre_hi = lam_hi >= 0 ? pre_lo + 3 : pre_lo;
        const int span = lam_hi >= 0 ? 2 : 1;
        if (G.compute_bins)
            for (int i = lane; i < span * nb * 4; i += 64) acc_lo[i] = 0.0;
        if (lane < 3 * span) pre_lo[lane] = 0.0;
        // the zeros are stored before any lane of this wave adds to them
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        wave_sync();
        double charge = 0.0;   // of this lane's half's wavelength
        bool any = G.opt.mus[lam_lo] != 0.0 || (lam_hi >= 0 && G.opt.mus[lam_hi] != 0.0);
        if (any) {
            for (int b = G.om_batches[w] + seg; b < G.om_batches[w + 1]; b += S) {
                BatchHead h = G.heads[b];
                if (h.n == 0) continue;
                size_t off = (size_t)G.node_off[b];
                charge += deposit_pair(G, lam_lo, lam_hi, h, G.pool_path + off, G.pool_g0 + off,
                                       G.pool_g2 + off, acc_lo, acc_hi, pre_lo, pre_hi, lane,
                                       ev_bin_all[wv], ev_val_all[wv], head_at_all[wv],
                                       count_panels, count_events);
                wave_sync();
            }
        }
        if (lane == 0) G.part_charge[slot * nl + lam_lo] = charge;
        if (lane == 32 && lam_hi >= 0) G.part_charge[slot * nl + lam_hi] = charge;
    }
    if (G.work_counts != nullptr) {
        if (lane == 0 && (count_panels || count_events)) {
            atomicAdd(&s_counts[0], count_panels);
            atomicAdd(&s_counts[1], count_events);
        }
        __syncthreads();
        if (tid == 0 && (s_counts[0] || s_counts[1])) {
            atomicAdd(G.work_counts + 1, s_counts[0]);
            atomicAdd(G.work_counts + 2, s_counts[1]);
        }
    }
}

// The single-kernel form: one workgroup of 8 waves per (OM, segment) builds a
// batch's path nodes and profile in LDS; then every wave deposits one
// wavelength. It runs when the node pool cannot hold every batch's path.
template <int P>
__global__ void __launch_bounds__(kBatchThreads) o1_batches(Group G, int groups) {
    if (wide_form(G, groups)) return;
    __shared__ double coeff[kBatchGroups][16];
    __shared__ double path[kMaxPathNodes];
    __shared__ double g0[kMaxPathNodes];
    __shared__ double g2[kMaxPathNodes];
    __shared__ int ev_bin_all[kBatchWaves][64 * kMaxDirect];
    __shared__ double ev_val_all[kBatchWaves][64 * kMaxDirect][4];
    __shared__ int head_at_all[kBatchWaves][64 * kMaxDirect];
    __shared__ double charge_acc[kMaxBands];
    __shared__ unsigned long long s_counts[3];
    int w = blockIdx.x, seg = blockIdx.y;
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int* ev_bin = ev_bin_all[wave];
    double (*ev_val)[4] = ev_val_all[wave];
    int* head_at = head_at_all[wave];
    int nl = G.opt.bands, nb = G.opt.bins;
    int nv = 2 * G.opt.n_mono + 6;
    size_t slot = (size_t)w * G.segments + seg;
    double* acc = G.part_acc + slot * nl * nb * 4;
    double* pre = G.part_pre + slot * nl * 3;
    if (G.compute_bins)
        for (size_t i = tid; i < (size_t)nl * nb * 4; i += kBatchThreads) acc[i] = 0.0;
    for (int i = tid; i < nl * 3; i += kBatchThreads) pre[i] = 0.0;
    if (tid < nl) charge_acc[tid] = 0.0;
    if (tid < 3) s_counts[tid] = 0;
    __syncthreads();
    unsigned long long count_terms = 0, count_panels = 0, count_events = 0;
    double gg = 2.0 * G.opt.g, g1 = 1.0 + G.opt.g * G.opt.g;
    double norm = (1.0 - G.opt.g * G.opt.g) / (4.0 * M_PI);
    for (int b = G.om_batches[w] + seg; b < G.om_batches[w + 1]; b += G.segments) {
        BatchHead h = batch_head(G, b, lane);
        if (h.n == 0) continue;
        int n = h.n;
        count_terms += (unsigned long long)n * (unsigned long long)(h.stop - h.start);
        __syncthreads();   // the previous batch's deposits have read the path
        for (int k = tid; k < n; k += kBatchThreads) {
            path[k] = path_node(h, k);
            g0[k] = 0.0;
            g2[k] = 0.0;
        }
        // profile: every path node sums its batch's groups in key order
        for (int c0 = h.start; c0 < h.stop; c0 += kBatchGroups) {
            int chunk = min(kBatchGroups, h.stop - c0);
            __syncthreads();
            for (int q = tid; q < chunk; q += kBatchThreads)
                stage_group<P>(G, G.values + (size_t)(c0 + q) * nv, coeff[q]);
            __syncthreads();
            for (int k = tid; k < n; k += kBatchThreads) {
                double acc0 = g0[k], acc2 = g2[k];
                double s_k = path[k];
                for (int q = 0; q < chunk; ++q)
                    profile_add<P>(G, coeff[q], G.values + (size_t)(c0 + q) * nv, s_k, h.r_b,
                                   h.backward, gg, g1, norm, acc0, acc2);
                g0[k] = acc0;
                g2[k] = acc2;
            }
        }
        __syncthreads();
        for (int lam = wave; lam < nl; lam += kBatchWaves) {
            if (G.opt.mus[lam] == 0.0) continue;
            double total = deposit_lambda(G, lam, h, path, g0, g2, acc + (size_t)lam * nb * 4,
                                          pre + 3 * lam, lane, ev_bin, ev_val, head_at,
                                          count_panels, count_events);
            if (lane == 0) charge_acc[lam] += total;
            wave_sync();
        }
    }
    __syncthreads();
    if (tid < nl) G.part_charge[slot * nl + tid] = charge_acc[tid];
    if (G.work_counts != nullptr) {
        // profile terms are the same in every wave: count them once
        if (tid == 0) s_counts[0] = count_terms;
        if (lane == 0) {
            atomicAdd(&s_counts[1], count_panels);
            atomicAdd(&s_counts[2], count_events);
        }
        __syncthreads();
        if (tid == 0) {
            atomicAdd(G.work_counts, s_counts[0]);
            atomicAdd(G.work_counts + 1, s_counts[1]);
            atomicAdd(G.work_counts + 2, s_counts[2]);
        }
    }
}

// Partial deposits of the segments summed in segment order into segment 0.
__global__ void __launch_bounds__(kThreads) o1_sum_segments(Group G) {
    int w = blockIdx.y;
    int nl = G.opt.bands, nb = G.opt.bins, S = G.segments;
    size_t per = (size_t)nl * nb * 4;
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    double* acc0 = G.part_acc + (size_t)w * S * per;
    if (i < per) {
        double v = acc0[i];
        for (int s = 1; s < S; ++s) v += acc0[(size_t)s * per + i];
        acc0[i] = v;
    }
    if (i < (size_t)nl * 3) {
        double* pre0 = G.part_pre + (size_t)w * S * nl * 3;
        double v = pre0[i];
        for (int s = 1; s < S; ++s) v += pre0[(size_t)s * nl * 3 + i];
        pre0[i] = v;
    }
}

// The per-wavelength prefix pass over the summed deposits (_finalize
// :610-632), one wave per (OM, wavelength): the running sums (a, b, c) at every
// bin are the window-front sums plus an exclusive scan of the bins' deposits,
// 64 bins at a time with the carry between chunks.
__global__ void __launch_bounds__(kThreads) o1_finalize_bins(Group G) {
    int lane = threadIdx.x & 63;
    long long gw = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    int nl = G.opt.bands, nb = G.opt.bins, S = G.segments;
    if (gw >= (long long)G.om.work_count * nl) return;
    int w = (int)(gw / nl), lam = (int)(gw % nl);
    size_t base = (size_t)w * S;
    const double* a4 = G.part_acc + base * nl * nb * 4 + (size_t)lam * nb * 4;
    const double* pre0 = G.part_pre + base * nl * 3;
    const double* eedge = G.opt.eedge + (size_t)lam * (nb + 1);
    const double* it = G.opt.itab + (size_t)lam * nb * 3;
    double* lb = G.lambda_bins + ((size_t)w * nl + lam) * nb;
    double a = pre0[3 * lam], b = pre0[3 * lam + 1], c = pre0[3 * lam + 2];
    for (int k0 = 0; k0 < nb; k0 += 64) {
        int k = k0 + lane;
        bool in = k < nb;
        double da = in ? a4[4 * (size_t)k] : 0.0;
        double db = in ? a4[4 * (size_t)k + 1] : 0.0;
        double dc = in ? a4[4 * (size_t)k + 2] : 0.0;
        double ia = da, ib = db, ic = dc;   // inclusive scans
        for (int off = 1; off < 64; off <<= 1) {
            double xa = __shfl_up(ia, off, 64), xb = __shfl_up(ib, off, 64),
                   xc = __shfl_up(ic, off, 64);
            if (lane >= off) { ia += xa; ib += xb; ic += xc; }
        }
        double ea = __shfl_up(ia, 1, 64), eb = __shfl_up(ib, 1, 64), ec = __shfl_up(ic, 1, 64);
        if (lane == 0) { ea = 0.0; eb = 0.0; ec = 0.0; }
        if (in) {
            double ak = a + ea, bk = b + eb, ck = c + ec;
            double e_k = G.opt.edges[k];
            double width = G.opt.edges[k + 1] - e_k;
            double a_loc = ak + e_k * (bk + ck * e_k);
            double b_loc = bk + 2.0 * ck * e_k;
            double value = eedge[k] * width
                           * (a_loc * it[3 * k] + width * (b_loc * it[3 * k + 1]
                                                           + width * ck * it[3 * k + 2]));
            lb[k] = value + a4[4 * (size_t)k + 3];
        }
        a += __shfl(ia, 63, 64);
        b += __shfl(ib, 63, 64);
        c += __shfl(ic, 63, 64);
    }
}

// Charge of every OM (segments, then wavelengths, in order) and its bins, the
// wavelengths' bins summed in order.
__global__ void __launch_bounds__(kThreads) o1_finalize(Group G) {
    int w = blockIdx.x;
    int d = G.om.work[w];
    int nl = G.opt.bands, nb = G.opt.bins, S = G.segments;
    size_t base = (size_t)w * S;
    if (threadIdx.x == 0) {
        double total = 0.0;
        for (int l = 0; l < nl; ++l) {
            double c = 0.0;
            for (int s = 0; s < S; ++s) c += G.part_charge[(base + s) * nl + l];
            total += c;
        }
        G.charge[d] = G.om.areas[d] * total;
    }
    if (!G.compute_bins) return;
    const double* lb = G.lambda_bins + (size_t)w * nl * nb;
    double area = G.om.areas[d];
    for (int k = threadIdx.x; k < nb; k += kThreads) {
        double total = 0.0;
        for (int l = 0; l < nl; ++l) total += area * lb[(size_t)l * nb + k];
        G.bins[(size_t)d * nb + k] = total;
    }
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int bits_for(unsigned long long value) {
    int bits = 1;
    while (bits < 63 && (1ULL << bits) <= value) ++bits;
    return bits;
}

// ---- compress layout

struct CompressLayout {
    size_t keys, ids, sorted_keys, sorted_ids, sub_element, sub_index, run_flag, run_scan;
    size_t run_start, run_key, rows, info, sort_temp, scan_temp, pkeys, pids, spkeys, spids,
        sub_record,
        pflag, pscan, psort_temp, piece_count, piece_start, big_count, big_start, piece_rows,
        piece_run, total;
};

int compress_layout(const lh_order1_grid* grid, int count, long long runs_cap,
                    CompressLayout& L) {
    long long nodes = grid->subsegments * grid->ring_nodes;
    size_t sort_temp = 0, scan_temp = 0, psort_temp = 0;
    unsigned long long max_key = (unsigned long long)grid->dims[0] * grid->dims[1] * grid->dims[2]
                                 * 6ULL * grid->pixel_face * grid->pixel_face;
    size_t sort_temp32 = 0;
    if (lh::sort_pairs(nullptr, sort_temp, (long long*)nullptr,
                                           (long long*)nullptr, (int*)nullptr, (int*)nullptr,
                                           (int)nodes, 0, bits_for(max_key), (hipStream_t)0)
            != hipSuccess ||
        lh::sort_pairs(nullptr, sort_temp32, (unsigned*)nullptr,
                                           (unsigned*)nullptr, (int*)nullptr, (int*)nullptr,
                                           (int)nodes, 0, 32, (hipStream_t)0) != hipSuccess) {
        lh::set_error("radix sort workspace query failed");
        return LH_ERR_HIP;
    }
    sort_temp = std::max(sort_temp, sort_temp32);
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_temp, (int*)nullptr, (int*)nullptr,
                                         (int)(nodes + 1), (hipStream_t)0) != hipSuccess) {
        lh::set_error("scan workspace query failed");
        return LH_ERR_HIP;
    }
    if (lh::sort_pairs(nullptr, psort_temp, (long long*)nullptr,
                                           (long long*)nullptr, (int*)nullptr, (int*)nullptr,
                                           (int)runs_cap, 0, bits_for(max_key),
                                           (hipStream_t)0) != hipSuccess) {
        lh::set_error("radix sort workspace query failed");
        return LH_ERR_HIP;
    }
    size_t at = 0;
    auto take = [&](size_t bytes) { size_t here = at; at += align_up(bytes); return here; };
    L.info = take(8 * sizeof(int));
    L.keys = take(nodes * sizeof(long long));
    L.ids = take(nodes * sizeof(int));
    L.sorted_keys = take(nodes * sizeof(long long));
    L.sorted_ids = take(nodes * sizeof(int));
    L.sub_element = take(grid->subsegments * sizeof(int));
    L.sub_index = take(grid->subsegments * sizeof(int));
    L.sub_record = take((size_t)grid->subsegments * kSubRecord * sizeof(double));
    L.run_flag = take((nodes + 1) * sizeof(int));
    L.run_scan = take((nodes + 1) * sizeof(int));
    L.run_start = take((nodes + 1) * sizeof(int));
    L.run_key = take(nodes * sizeof(long long));
    L.rows = take((size_t)runs_cap * kNmom * sizeof(double));
    L.pkeys = take(runs_cap * sizeof(long long));
    L.pids = take(runs_cap * sizeof(int));
    L.spkeys = take(runs_cap * sizeof(long long));
    L.spids = take(runs_cap * sizeof(int));
    L.pflag = take((runs_cap + 1) * sizeof(int));
    L.pscan = take((runs_cap + 1) * sizeof(int));
    L.piece_count = take((runs_cap + 1) * sizeof(int));
    L.piece_start = take((runs_cap + 1) * sizeof(int));
    L.big_count = take((runs_cap + 1) * sizeof(int));
    L.big_start = take((runs_cap + 1) * sizeof(int));
    // pieces of runs longer than one piece: at most 2 nodes / kRowPiece
    L.piece_rows = take((size_t)(2 * nodes / kRowPiece + 1) * kNmom * sizeof(double));
    // pieces <= runs + nodes / kRowPiece
    L.piece_run = take((size_t)(runs_cap + nodes / kRowPiece + 1) * sizeof(int));
    L.sort_temp = take(sort_temp);
    L.scan_temp = take(scan_temp);
    L.psort_temp = take(psort_temp);
    L.total = at;
    return LH_OK;
}

int check_grid(const lh_order1_chords* ch, const lh_order1_grid* grid) {
    LH_REQUIRE(ch != nullptr && grid != nullptr, "descriptor is NULL");
    LH_REQUIRE(ch->count >= 1, "no chords");
    LH_REQUIRE(ch->start_m && ch->direction && ch->length_m && ch->coefficient0
                   && ch->coefficient2 && ch->cone_cosine && ch->start_ns && ch->end_ns
                   && ch->sub_offset, "chord pointers must be set");
    LH_REQUIRE(grid->cell_m > 0 && grid->pixel_face >= 1 && grid->ring_nodes >= 1,
               "cell_m, pixel_face and ring_nodes must be positive");
    LH_REQUIRE(grid->subsegments >= ch->count, "subsegment count is inconsistent");
    long long nodes = grid->subsegments * grid->ring_nodes;
    LH_REQUIRE(nodes < (1LL << 31) - 1, "too many ring nodes (%lld)", nodes);
    double cells = (double)grid->dims[0] * grid->dims[1] * grid->dims[2];
    LH_REQUIRE(grid->dims[0] >= 1 && grid->dims[1] >= 1 && grid->dims[2] >= 1
                   && cells * 6.0 * grid->pixel_face * grid->pixel_face < 4.0e18,
               "cell grid too large for 64-bit pencil keys; increase cell_m");
    LH_REQUIRE(grid->parent_face == 0
                   || (grid->parent_face > 0 && grid->pixel_face % grid->parent_face == 0),
               "angular refinement must be a multiple of its parent");
    return LH_OK;
}

Compress make_compress(const lh_order1_chords* ch, const lh_order1_grid* grid, char* base,
                       const CompressLayout& L) {
    Compress k;
    k.ch = *ch;
    k.grid = *grid;
    k.nodes = grid->subsegments * grid->ring_nodes;
    k.keys = reinterpret_cast<long long*>(base + L.keys);
    k.ids = reinterpret_cast<int*>(base + L.ids);
    k.sorted_keys = reinterpret_cast<long long*>(base + L.sorted_keys);
    k.sorted_ids = reinterpret_cast<int*>(base + L.sorted_ids);
    k.sub_element = reinterpret_cast<int*>(base + L.sub_element);
    k.sub_index = reinterpret_cast<int*>(base + L.sub_index);
    k.run_flag = reinterpret_cast<int*>(base + L.run_flag);
    k.run_scan = reinterpret_cast<int*>(base + L.run_scan);
    k.run_start = reinterpret_cast<int*>(base + L.run_start);
    k.run_key = reinterpret_cast<long long*>(base + L.run_key);
    k.sub_record = reinterpret_cast<double*>(base + L.sub_record);
    const unsigned long long max_key = (unsigned long long)grid->dims[0] * grid->dims[1]
                                       * grid->dims[2] * 6ULL * grid->pixel_face
                                       * grid->pixel_face;
    k.key32 = bits_for(max_key) <= 32;
    k.rows = reinterpret_cast<double*>(base + L.rows);
    k.info = reinterpret_cast<int*>(base + L.info);
    return k;
}

unsigned blocks(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// ---- grouped layout

struct GroupLayout {
    size_t keys, vals, sorted_keys, sorted_vals, group_flag, group_scan, group_start, group_end,
        om_groups, prepared, lambda_bins, k_range, x_range, max_key, info, sort_temp, scan_temp,
        total;
};

// Entries per (pencil, OM): the screen's radial rule may give two nodes.
inline int slots_of(const lh_order1_pencils* pen) { return pen->parent != nullptr ? 2 : 1; }

int grouped_layout(int work, int n_pen, int slots, int bands, int bins, GroupLayout& L) {
    long long entries = (long long)slots * work * n_pen;
    // entries, groups and their offsets are 32-bit on the device
    LH_REQUIRE(entries < (1LL << 31) - 1,
               "%lld (pencil, OM) entries in one grouped call; pass fewer OMs per call",
               entries);
    size_t sort_temp = 0, sort_temp32 = 0, scan_temp = 0;
    if (lh::sort_pairs(
            nullptr, sort_temp, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
            (int*)nullptr, (int*)nullptr, (int)entries, 0, 64, (hipStream_t)0) != hipSuccess ||
        lh::sort_pairs(
            nullptr, sort_temp32, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
            (int*)nullptr, (int)entries, 0, 32, (hipStream_t)0) != hipSuccess) {
        lh::set_error("sort workspace query failed");
        return LH_ERR_HIP;
    }
    sort_temp = std::max(sort_temp, sort_temp32);
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_temp, (int*)nullptr, (int*)nullptr,
                                         (int)(entries + 1), (hipStream_t)0) != hipSuccess) {
        lh::set_error("scan workspace query failed");
        return LH_ERR_HIP;
    }
    size_t at = 0;
    auto take = [&](size_t bytes) { size_t here = at; at += align_up(bytes); return here; };
    L.info = take(8 * sizeof(int));
    L.max_key = take(sizeof(unsigned long long));
    L.k_range = take((size_t)work * 4 * sizeof(int));
    L.x_range = take((size_t)work * 4 * sizeof(unsigned long long));
    L.keys = take(entries * sizeof(long long));
    L.vals = take(entries * sizeof(int));
    L.sorted_keys = take(entries * sizeof(long long));
    L.sorted_vals = take(entries * sizeof(int));
    L.group_flag = take((entries + 1) * sizeof(int));
    L.group_scan = take((entries + 1) * sizeof(int));
    L.group_start = take((entries + 1) * sizeof(int));
    L.group_end = take((entries + 1) * sizeof(int));
    L.om_groups = take((work + 1) * sizeof(int));
    L.prepared = take((size_t)n_pen * 7 * sizeof(double));
    L.lambda_bins = take((size_t)work * bands * bins * sizeof(double));
    L.sort_temp = take(sort_temp);
    L.scan_temp = take(scan_temp);
    L.total = at;
    return LH_OK;
}

// Batch segments per OM: about 8192 deposit waves (OM x segment x wavelength)
// in all, within 2 GiB of per-slot partial deposits.
int segments_for(int work, int bands, int bins) {
    long long lanes = (long long)work * std::max(bands, 1);
    int S = (int)std::max(1LL, std::min(64LL, (8192 + lanes - 1) / lanes));
    while (S > 1 && (double)work * S * bands * bins * 32.0 > 2048.0 * 1024 * 1024) S /= 2;
    return S;
}

struct ScratchLayout {
    size_t values, batch_flag, batch_scan, batch_start, om_batches, part_acc, part_pre,
        part_charge, piece_count, member_off, piece_start, tile_group, piece_values, heads, node_n,
        node_off, scan_temp, total;   // total: the fixed part; the node pool follows it
    int segments;
    long long pieces;         // bound on the group-sum pieces
};

int scratch_layout(int work, int n_pen, int slots, int bands, int bins, int n_mono, int groups,
                   ScratchLayout& L) {
    size_t scan_temp = 0, scan_nodes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_temp, (int*)nullptr, (int*)nullptr,
                                         groups + 1, (hipStream_t)0) != hipSuccess ||
        hipcub::DeviceScan::ExclusiveSum(nullptr, scan_nodes, (long long*)nullptr,
                                         (long long*)nullptr, groups + 1,
                                         (hipStream_t)0) != hipSuccess) {
        lh::set_error("scan workspace query failed");
        return LH_ERR_HIP;
    }
    scan_temp = std::max(scan_temp, scan_nodes);
    L.segments = segments_for(work, bands, bins);
    size_t parts = (size_t)work * L.segments;
    size_t at = 0;
    auto take = [&](size_t bytes) { size_t here = at; at += align_up(bytes); return here; };
    L.values = take((size_t)std::max(groups, 1) * (2 * n_mono + 6) * sizeof(double));
    L.batch_flag = take((groups + 1) * sizeof(int));
    L.batch_scan = take((groups + 1) * sizeof(int));
    L.batch_start = take((groups + 1) * sizeof(int));
    L.om_batches = take((work + 1) * sizeof(int));
    L.part_acc = take(parts * bands * bins * 4 * sizeof(double));
    L.part_pre = take(parts * bands * 3 * sizeof(double));
    L.part_charge = take(parts * bands * sizeof(double));
    long long entries = (long long)slots * work * n_pen;
    L.pieces = groups + (entries + kPieceEntries - 1) / kPieceEntries;
    L.piece_count = take((groups + 1) * sizeof(int));
    L.member_off = take((groups + 1) * sizeof(int));
    L.piece_start = take((groups + 1) * sizeof(int));
    L.tile_group = take((size_t)((entries + kPieceEntries - 1) / kPieceEntries + 1) * sizeof(int));
    L.piece_values = take((size_t)std::max(L.pieces, 1LL) * (2 * n_mono + 6) * sizeof(double));
    L.heads = take((size_t)(groups + 1) * sizeof(BatchHead));
    L.node_n = take((size_t)(groups + 1) * sizeof(long long));
    L.node_off = take((size_t)(groups + 1) * sizeof(long long));
    L.scan_temp = take(scan_temp);
    L.total = at;
    return LH_OK;
}

// Path nodes the pool of a scratch buffer of `bytes` holds beyond the fixed part
// (three FP64 arrays, each 256-byte aligned).
long long pool_nodes_of(const ScratchLayout& L, size_t bytes) {
    if (bytes <= L.total + 3 * 256) return 0;
    return (long long)((bytes - L.total - 3 * 256) / (3 * sizeof(double)));
}

// The default pool: 64 nodes per group and at least 2^20 (path grids have
// tens of nodes; a batch holds one group or more).
long long default_pool_nodes(int groups) {
    return std::max(64LL * groups, 1LL << 20);
}

int check_grouped(const lh_order1_pencils* pen, const lh_order1_oms* om,
                  const lh_order1_optics* opt, const lh_order1_merge* merge) {
    LH_REQUIRE(pen && om && opt && merge, "descriptor is NULL");
    LH_REQUIRE(pen->count >= 1 && om->work_count >= 1, "no pencils or no OMs");
    LH_REQUIRE(om->work_count <= 65535, "at most 65535 OMs per grouped call");
    LH_REQUIRE(opt->bands >= 1 && opt->bands <= kMaxBands, "1..%d wavelength bands", kMaxBands);
    LH_REQUIRE(opt->max_power >= 0 && opt->max_power <= 4
                   && opt->n_mono == (opt->max_power + 1) * (opt->max_power + 2) / 2,
               "acceptance degree must lie in 0..4 with (P+1)(P+2)/2 monomials");
    LH_REQUIRE(opt->bins >= 1 && opt->edges && opt->eedge && opt->itab, "bins and tables needed");
    LH_REQUIRE(opt->direct_bins >= 0 && opt->direct_bins <= kMaxDirect,
               "direct_bins must lie in 0..%d", kMaxDirect);
    LH_REQUIRE(opt->max_path_nodes >= 3 && opt->max_path_nodes <= kMaxPathNodes,
               "max_path_nodes must lie in 3..%d", kMaxPathNodes);
    LH_REQUIRE(2LL * om->work_count * pen->count < (1LL << 31) - 1,
               "too many (OM, pencil) entries; split the OM list");
    LH_REQUIRE(merge->log_r_step >= 710.0 / ((1 << 20) - 2),
               "grouped radius/angular resolution exceeds integer key range");
    return LH_OK;
}

Group make_group(const lh_order1_pencils* pen, const lh_order1_oms* om,
                 const lh_order1_optics* opt, const lh_order1_merge* merge, char* base,
                 const GroupLayout& L) {
    Group G;
    memset(&G, 0, sizeof(G));
    G.pen = *pen;
    G.om = *om;
    G.opt = *opt;
    G.merge = *merge;
    G.screen = pen->parent != nullptr;
    G.slots = slots_of(pen);
    G.n_pen = pen->count;
    G.log_ratio = log(1.0 + merge->theta_ratio);
    G.theta_bins = 2 + (long long)(log(M_PI / merge->theta_lo) / G.log_ratio);
    G.keys = reinterpret_cast<long long*>(base + L.keys);
    G.vals = reinterpret_cast<int*>(base + L.vals);
    G.sorted_keys = reinterpret_cast<long long*>(base + L.sorted_keys);
    G.sorted_vals = reinterpret_cast<int*>(base + L.sorted_vals);
    G.group_flag = reinterpret_cast<int*>(base + L.group_flag);
    G.group_scan = reinterpret_cast<int*>(base + L.group_scan);
    G.group_start = reinterpret_cast<int*>(base + L.group_start);
    G.group_end = reinterpret_cast<int*>(base + L.group_end);
    G.om_groups = reinterpret_cast<int*>(base + L.om_groups);
    G.prepared = reinterpret_cast<double*>(base + L.prepared);
    G.lambda_bins = reinterpret_cast<double*>(base + L.lambda_bins);
    G.k_range = reinterpret_cast<int*>(base + L.k_range);
    G.x_range = reinterpret_cast<unsigned long long*>(base + L.x_range);
    G.max_key = reinterpret_cast<unsigned long long*>(base + L.max_key);
    G.info = reinterpret_cast<int*>(base + L.info);
    G.mut_min = opt->mut_min;
    G.speed_max = opt->speed_max;
    return G;
}

}  // namespace

extern "C" {

int lh_order1_set_tables(const double* taylor, const double* tangent,
                         const double* weighted_sec2) {
    LH_REQUIRE(taylor && tangent && weighted_sec2, "tables must be set");
    LH_HIP_CALL(hipMemcpyToSymbol(HIP_SYMBOL(kTaylor), taylor, sizeof(double) * 72));
    LH_HIP_CALL(hipMemcpyToSymbol(HIP_SYMBOL(kTangent), tangent, sizeof(double) * kGauss));
    LH_HIP_CALL(hipMemcpyToSymbol(HIP_SYMBOL(kWeightedSec2), weighted_sec2,
                                  sizeof(double) * kGauss));
    return LH_OK;
}

size_t lh_order1_compress_workspace_bytes(const lh_order1_chords* chords,
                                          const lh_order1_grid* grid) {
    if (check_grid(chords, grid) != LH_OK) return 0;
    CompressLayout L;
    long long nodes = grid->subsegments * grid->ring_nodes;
    if (compress_layout(grid, chords->count, nodes, L) != LH_OK) return 0;
    return L.total;
}

int lh_order1_compress_keys(const lh_order1_chords* chords, const lh_order1_grid* grid, int* info,
                            void* workspace, size_t workspace_bytes, void* stream) {
    int status = check_grid(chords, grid);
    if (status != LH_OK) return status;
    CompressLayout L;
    long long nodes = grid->subsegments * grid->ring_nodes;
    if ((status = compress_layout(grid, chords->count, nodes, L)) != LH_OK) return status;
    LH_REQUIRE(workspace && workspace_bytes >= L.total && info,
               "workspace too small: %zu bytes given, %zu needed", workspace_bytes, L.total);
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    Compress k = make_compress(chords, grid, base, L);
    o1_subsegments<<<blocks(chords->count), kThreads, 0, s>>>(k);
    if ((status = lh::check_launch("o1_subsegments")) != LH_OK) return status;
    // the per-sub-segment record, read by o1_nodes here and by the moment
    // rows of lh_order1_compress_pencils (same workspace)
    o1_sub_records<<<blocks(grid->subsegments), kThreads, 0, s>>>(k);
    if ((status = lh::check_launch("o1_sub_records")) != LH_OK) return status;
    o1_nodes<<<(unsigned)((grid->subsegments * (long long)grid->ring_nodes + kThreads - 1)
                          / kThreads), kThreads, 0, s>>>(k);
    if ((status = lh::check_launch("o1_nodes")) != LH_OK) return status;
    unsigned long long max_key = (unsigned long long)grid->dims[0] * grid->dims[1] * grid->dims[2]
                                 * 6ULL * grid->pixel_face * grid->pixel_face;
    size_t sort_bytes = L.scan_temp - L.sort_temp;
    if (k.key32) {
        LH_HIP_CALL(lh::sort_pairs(
            base + L.sort_temp, sort_bytes, reinterpret_cast<unsigned*>(k.keys),
            reinterpret_cast<unsigned*>(k.sorted_keys), k.ids, k.sorted_ids, (int)nodes, 0,
            bits_for(max_key), s));
    } else {
        LH_HIP_CALL(lh::sort_pairs(base + L.sort_temp, sort_bytes, k.keys,
                                                       k.sorted_keys, k.ids, k.sorted_ids,
                                                       (int)nodes, 0, bits_for(max_key), s));
    }
    o1_run_flags<<<blocks(nodes + 1), kThreads, 0, s>>>(k);
    if ((status = lh::check_launch("o1_run_flags")) != LH_OK) return status;
    size_t scan_bytes = L.psort_temp - L.scan_temp;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + L.scan_temp, scan_bytes, k.run_flag,
                                                 k.run_scan, (int)(nodes + 1), s));
    o1_run_starts<<<blocks(nodes + 1), kThreads, 0, s>>>(k);
    if ((status = lh::check_launch("o1_run_starts")) != LH_OK) return status;
    LH_HIP_CALL(hipMemcpyAsync(info, k.info, sizeof(int), hipMemcpyDeviceToDevice, s));
    return LH_OK;
}

int lh_order1_compress_pencils(const lh_order1_chords* chords, const lh_order1_grid* grid,
                               int runs, lh_order1_pencils* out, int* info, void* workspace,
                               size_t workspace_bytes, void* stream) {
    int status = check_grid(chords, grid);
    if (status != LH_OK) return status;
    CompressLayout L;
    long long nodes = grid->subsegments * grid->ring_nodes;
    if ((status = compress_layout(grid, chords->count, nodes, L)) != LH_OK) return status;
    LH_REQUIRE(workspace && workspace_bytes >= L.total, "workspace too small");
    LH_REQUIRE(out && runs >= 1 && runs <= nodes, "invalid pencil count");
    LH_REQUIRE(out->position && out->direction && out->w0 && out->w2 && out->time && out->major
                   && out->var_major && out->var_minor && out->pos_cov, "pencil arrays must be set");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    Compress k = make_compress(chords, grid, base, L);
    {
        int* counts = reinterpret_cast<int*>(base + L.piece_count);
        int* piece_start = reinterpret_cast<int*>(base + L.piece_start);
        int* big = reinterpret_cast<int*>(base + L.big_count);
        int* big_start = reinterpret_cast<int*>(base + L.big_start);
        double* piece_rows = reinterpret_cast<double*>(base + L.piece_rows);
        int* piece_run = reinterpret_cast<int*>(base + L.piece_run);
        o1_row_counts<<<blocks(runs + 1), kThreads, 0, s>>>(k, runs, counts, big);
        if ((status = lh::check_launch("o1_row_counts")) != LH_OK) return status;
        size_t scan_bytes = L.psort_temp - L.scan_temp;
        LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + L.scan_temp, scan_bytes, counts,
                                                     piece_start, runs + 1, s));
        LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + L.scan_temp, scan_bytes, big,
                                                     big_start, runs + 1, s));
        // pieces <= runs + nodes / kRowPiece
        long long bound = runs + nodes / kRowPiece + 1;
        o1_piece_runs<<<blocks(runs), kThreads, 0, s>>>(runs, piece_start, piece_run);
        if ((status = lh::check_launch("o1_piece_runs")) != LH_OK) return status;
        {
            lh::KernelMark mark("o1_row_pieces", s);
            o1_row_pieces<<<(unsigned)((bound + 3) / 4), kThreads, 0, s>>>(k, runs, piece_start,
                                                                          big_start, piece_run,
                                                                          piece_rows);
        }
        if ((status = lh::check_launch("o1_row_pieces")) != LH_OK) return status;
        o1_row_combine<<<blocks(runs), kThreads, 0, s>>>(k, runs, piece_start, big_start,
                                                         piece_rows);
        if ((status = lh::check_launch("o1_row_combine")) != LH_OK) return status;
    }
    if (grid->parent_face == 0) {
        LH_REQUIRE(out->count >= runs, "pencil capacity %d below %d", out->count, runs);
        o1_finish<<<blocks(runs), kThreads, 0, s>>>(k.rows, runs, *out);
        return lh::check_launch("o1_finish");
    }
    LH_REQUIRE(out->parent && out->next && info, "screen pencils need parent, next and info");
    LH_REQUIRE(out->count >= 2 * runs, "screen pencil capacity %d below %d", out->count, 2 * runs);
    long long* pkeys = reinterpret_cast<long long*>(base + L.pkeys);
    int* pids = reinterpret_cast<int*>(base + L.pids);
    long long* spkeys = reinterpret_cast<long long*>(base + L.spkeys);
    int* spids = reinterpret_cast<int*>(base + L.spids);
    int* pflag = reinterpret_cast<int*>(base + L.pflag);
    int* pscan = reinterpret_cast<int*>(base + L.pscan);
    o1_parent_keys<<<blocks(runs), kThreads, 0, s>>>(k, runs, pkeys, pids);
    if ((status = lh::check_launch("o1_parent_keys")) != LH_OK) return status;
    unsigned long long max_key = (unsigned long long)grid->dims[0] * grid->dims[1] * grid->dims[2]
                                 * 6ULL * grid->pixel_face * grid->pixel_face;
    size_t psort_bytes = L.total - L.psort_temp;
    // pencils are in fine-key order: a stable sort by parent key gives (parent, fine) order
    LH_HIP_CALL(lh::sort_pairs(base + L.psort_temp, psort_bytes, pkeys,
                                                   spkeys, pids, spids, runs, 0,
                                                   bits_for(max_key), s));
    o1_parent_flags<<<blocks(runs + 1), kThreads, 0, s>>>(spkeys, runs, pflag);
    if ((status = lh::check_launch("o1_parent_flags")) != LH_OK) return status;
    size_t scan_bytes = L.psort_temp - L.scan_temp;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + L.scan_temp, scan_bytes, pflag, pscan,
                                                 runs + 1, s));
    o1_parents<<<blocks(runs), kThreads, 0, s>>>(k.rows, spids, pflag, pscan, runs, *out);
    if ((status = lh::check_launch("o1_parents")) != LH_OK) return status;
    o1_parent_links<<<blocks(runs), kThreads, 0, s>>>(pflag, pscan, runs, *out);
    if ((status = lh::check_launch("o1_parent_links")) != LH_OK) return status;
    LH_HIP_CALL(hipMemcpyAsync(info + 1, pscan + runs, sizeof(int), hipMemcpyDeviceToDevice, s));
    return LH_OK;
}

size_t lh_order1_grouped_workspace_bytes(const lh_order1_pencils* pencils,
                                         const lh_order1_oms* oms,
                                         const lh_order1_optics* optics) {
    if (!pencils || !oms || !optics) return 0;
    GroupLayout L;
    if (grouped_layout(oms->work_count, pencils->count, slots_of(pencils), optics->bands,
                       optics->bins, L) != LH_OK)
        return 0;
    return L.total;
}

size_t lh_order1_group_scratch_bytes(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                                     const lh_order1_optics* optics, int groups,
                                     long long pool_nodes) {
    if (!pencils || !oms || !optics || groups < 0) return 0;
    ScratchLayout L;
    if (scratch_layout(oms->work_count, pencils->count, slots_of(pencils), optics->bands,
                       optics->bins, optics->n_mono, groups, L) != LH_OK) return 0;
    long long pool = pool_nodes < 0 ? default_pool_nodes(groups) : pool_nodes;
    return pool == 0 ? L.total : L.total + 3 * 256 + (size_t)pool * 3 * sizeof(double);
}

int lh_order1_group_keys(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                         const lh_order1_optics* optics, const lh_order1_merge* merge, int* info,
                         void* workspace, size_t workspace_bytes, void* stream) {
    int status = check_grouped(pencils, oms, optics, merge);
    if (status != LH_OK) return status;
    GroupLayout L;
    if ((status = grouped_layout(oms->work_count, pencils->count, slots_of(pencils),
                                 optics->bands, optics->bins, L)) != LH_OK) return status;
    LH_REQUIRE(workspace && workspace_bytes >= L.total && info, "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    Group G = make_group(pencils, oms, optics, merge, base, L);
    int work = oms->work_count;
    long long total = (long long)G.slots * work * pencils->count;
    o1_prepare_shape<<<blocks(pencils->count), kThreads, 0, s>>>(G);
    if ((status = lh::check_launch("o1_prepare_shape")) != LH_OK) return status;
    o1_range_init<<<blocks(work), kThreads, 0, s>>>(G, work);
    if ((status = lh::check_launch("o1_range_init")) != LH_OK) return status;
    dim3 range_grid(std::min(blocks(pencils->count), (unsigned)kRangeGroups), work);
    o1_group_ranges<<<range_grid, kThreads, 0, s>>>(G);
    dim3 grid(blocks(pencils->count), work);
    if ((status = lh::check_launch("o1_group_ranges")) != LH_OK) return status;
    o1_key_space<<<blocks(work), kThreads, 0, s>>>(G, work);
    if ((status = lh::check_launch("o1_key_space")) != LH_OK) return status;
    // keys hold (OM << key_bits) | compact key; the compact sentinel (all ones)
    // lies above every compact key of its OM, so one sort over all OMs orders
    // every OM's entries by key with the skipped ones last
    unsigned long long space = 0;
    LH_HIP_CALL(hipMemcpyAsync(&space, G.max_key, sizeof(space), hipMemcpyDeviceToHost, s));
    LH_HIP_CALL(hipStreamSynchronize(s));
    G.key_bits = bits_for(space);
    G.key_mask = (1LL << G.key_bits) - 1;
    int end_bit = G.key_bits + bits_for((unsigned long long)std::max(work - 1, 0));
    LH_REQUIRE(end_bit <= 63, "grouped keys need %d bits; split the OM list", end_bit);
    // keys of at most 32 bits are stored and sorted as uint32 (a third less
    // data per sort pass)
    G.key32 = end_bit <= 32;
    {
        lh::KernelMark mark("o1_group_keys", s);
        o1_group_keys<<<grid, kThreads, 0, s>>>(G);
    }
    if ((status = lh::check_launch("o1_group_keys")) != LH_OK) return status;
    size_t sort_bytes = L.scan_temp - L.sort_temp;
    {
        lh::KernelMark mark("o1_sort", s);
        if (G.key32) {
            LH_HIP_CALL(lh::sort_pairs(
                base + L.sort_temp, sort_bytes, reinterpret_cast<unsigned*>(G.keys),
                reinterpret_cast<unsigned*>(G.sorted_keys), G.vals, G.sorted_vals, (int)total,
                0, end_bit, s));
        } else {
            LH_HIP_CALL(lh::sort_pairs(
                base + L.sort_temp, sort_bytes, reinterpret_cast<unsigned long long*>(G.keys),
                reinterpret_cast<unsigned long long*>(G.sorted_keys), G.vals, G.sorted_vals,
                (int)total, 0, end_bit, s));
        }
    }
    o1_group_flags<<<blocks(total + 1), kThreads, 0, s>>>(G, total);
    if ((status = lh::check_launch("o1_group_flags")) != LH_OK) return status;
    size_t scan_bytes = L.total - L.scan_temp;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(base + L.scan_temp, scan_bytes, G.group_flag,
                                                 G.group_scan, (int)(total + 1), s));
    o1_group_starts<<<blocks(total + 1), kThreads, 0, s>>>(G, total, work);
    if ((status = lh::check_launch("o1_group_starts")) != LH_OK) return status;
    LH_HIP_CALL(hipMemcpyAsync(info, G.info, sizeof(int), hipMemcpyDeviceToDevice, s));
    int host_bits = G.key_bits;
    LH_HIP_CALL(hipMemcpyAsync(info + 1, &host_bits, sizeof(int), hipMemcpyHostToDevice, s));
    LH_HIP_CALL(hipStreamSynchronize(s));
    return LH_OK;
}

int lh_order1_group_charge(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                           const lh_order1_optics* optics, const lh_order1_merge* merge,
                           int groups, int key_bits, void* scratch, size_t scratch_bytes,
                           int compute_bins, double* charge, double* bins,
                           unsigned long long* work_counts, void* workspace,
                           size_t workspace_bytes, void* stream) {
    int status = check_grouped(pencils, oms, optics, merge);
    if (status != LH_OK) return status;
    GroupLayout L;
    if ((status = grouped_layout(oms->work_count, pencils->count, slots_of(pencils),
                                 optics->bands, optics->bins, L)) != LH_OK) return status;
    LH_REQUIRE(workspace && workspace_bytes >= L.total, "workspace too small: %zu < %zu",
               workspace_bytes, L.total);
    LH_REQUIRE(charge != nullptr && (!compute_bins || bins != nullptr),
               "charge (and bins) must be set");
    LH_REQUIRE(optics->mut_min > 0 && optics->speed_max > 0, "mut_min and speed_max must be set");
    int work = oms->work_count;
    ScratchLayout SL;
    LH_REQUIRE(key_bits >= 1 && key_bits <= 62, "key_bits must come from lh_order1_group_keys");
    if ((status = scratch_layout(work, pencils->count, slots_of(pencils), optics->bands,
                                 optics->bins, optics->n_mono, std::max(groups, 0), SL)) != LH_OK)
        return status;
    LH_REQUIRE(groups >= 0 && scratch && scratch_bytes >= SL.total,
               "scratch too small: %zu < %zu", scratch_bytes, SL.total);
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    char* sc = static_cast<char*>(scratch);
    Group G = make_group(pencils, oms, optics, merge, base, L);
    G.key_bits = key_bits;
    G.key_mask = (1LL << key_bits) - 1;
    G.key32 = key_bits + bits_for((unsigned long long)std::max(work - 1, 0)) <= 32;
    G.compute_bins = compute_bins;
    G.work_counts = work_counts;
    G.charge = charge;
    G.bins = bins;
    G.values = reinterpret_cast<double*>(sc + SL.values);
    G.batch_flag = reinterpret_cast<int*>(sc + SL.batch_flag);
    G.batch_scan = reinterpret_cast<int*>(sc + SL.batch_scan);
    G.batch_start = reinterpret_cast<int*>(sc + SL.batch_start);
    G.om_batches = reinterpret_cast<int*>(sc + SL.om_batches);
    G.part_acc = reinterpret_cast<double*>(sc + SL.part_acc);
    G.part_pre = reinterpret_cast<double*>(sc + SL.part_pre);
    G.part_charge = reinterpret_cast<double*>(sc + SL.part_charge);
    G.segments = SL.segments;
    G.piece_start = reinterpret_cast<int*>(sc + SL.piece_start);
    G.member_off = reinterpret_cast<int*>(sc + SL.member_off);
    G.tile_group = reinterpret_cast<int*>(sc + SL.tile_group);
    G.piece_values = reinterpret_cast<double*>(sc + SL.piece_values);
    G.heads = reinterpret_cast<BatchHead*>(sc + SL.heads);
    G.node_n = reinterpret_cast<long long*>(sc + SL.node_n);
    G.node_off = reinterpret_cast<long long*>(sc + SL.node_off);
    G.pool_nodes = pool_nodes_of(SL, scratch_bytes);
    {
        size_t pool_bytes = align_up((size_t)G.pool_nodes * sizeof(double));
        G.pool_path = reinterpret_cast<double*>(sc + SL.total);
        G.pool_g0 = reinterpret_cast<double*>(sc + SL.total + pool_bytes);
        G.pool_g2 = reinterpret_cast<double*>(sc + SL.total + 2 * pool_bytes);
    }
    size_t scan_bytes = SL.total - SL.scan_temp;
    if (groups > 0) {
        int* counts = reinterpret_cast<int*>(sc + SL.piece_count);
        o1_member_counts<<<blocks(groups + 1), kThreads, 0, s>>>(G, groups, counts);
        if ((status = lh::check_launch("o1_member_counts")) != LH_OK) return status;
        LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(sc + SL.scan_temp, scan_bytes, counts,
                                                     G.member_off, groups + 1, s));
        o1_piece_counts<<<blocks(groups + 1), kThreads, 0, s>>>(G, groups, counts);
        if ((status = lh::check_launch("o1_piece_counts")) != LH_OK) return status;
        LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(sc + SL.scan_temp, scan_bytes, counts,
                                                     G.piece_start, groups + 1, s));
        o1_tile_groups<<<blocks(groups), kThreads, 0, s>>>(G, groups);
        if ((status = lh::check_launch("o1_tile_groups")) != LH_OK) return status;
        {
            lh::KernelMark mark("o1_group_sums", s);
            // one wave per tile of members; their count is on the device and
            // at most the entries' (waves past it return at once)
            long long entries = (long long)slots_of(pencils) * work * pencils->count;
            const unsigned grid_sums = (unsigned)((entries / kPieceEntries + 1 + 3) / 4);
            switch (optics->max_power) {
                case 0: o1_group_sums<0><<<grid_sums, kThreads, 0, s>>>(G, groups); break;
                case 1: o1_group_sums<1><<<grid_sums, kThreads, 0, s>>>(G, groups); break;
                case 2: o1_group_sums<2><<<grid_sums, kThreads, 0, s>>>(G, groups); break;
                case 3: o1_group_sums<3><<<grid_sums, kThreads, 0, s>>>(G, groups); break;
                default: o1_group_sums<4><<<grid_sums, kThreads, 0, s>>>(G, groups); break;
            }
        }
        if ((status = lh::check_launch("o1_group_sums")) != LH_OK) return status;
        // one thread per (tile, value); the tiles are at most the entries / 64 + 1
        long long entries = (long long)slots_of(pencils) * work * pencils->count;
        o1_group_combine<<<(unsigned)(((entries / kPieceEntries + 1) * (2 * optics->n_mono + 6)
                                       + kThreads - 1) / kThreads),
                           kThreads, 0, s>>>(G, groups);
        if ((status = lh::check_launch("o1_group_combine")) != LH_OK) return status;
    }
    o1_batch_flags<<<blocks(groups + 1), kThreads, 0, s>>>(G, groups, work);
    if ((status = lh::check_launch("o1_batch_flags")) != LH_OK) return status;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(sc + SL.scan_temp, scan_bytes, G.batch_flag,
                                                 G.batch_scan, groups + 1, s));
    o1_batch_starts<<<blocks(std::max(groups, work) + 1), kThreads, 0, s>>>(G, groups, work);
    if ((status = lh::check_launch("o1_batch_starts")) != LH_OK) return status;
    // batch headers and node counts, the node offsets, then either the wide
    // kernels or (the pool too small) the single-kernel form; each kernel
    // reads the total from the device and the other form's kernels return
    const unsigned grid_batches = (unsigned)std::min<long long>(
        ((long long)groups + 1 + kWideWaves - 1) / kWideWaves, kWideMaxBlocks);
    o1_batch_heads<<<grid_batches, kWideThreads, 0, s>>>(G, groups);
    if ((status = lh::check_launch("o1_batch_heads")) != LH_OK) return status;
    LH_HIP_CALL(hipcub::DeviceScan::ExclusiveSum(sc + SL.scan_temp, scan_bytes, G.node_n,
                                                 G.node_off, groups + 1, s));
    {
        lh::KernelMark mark("o1_batch_profile", s);
        switch (optics->max_power) {
            case 0: o1_batch_profile<0><<<grid_batches, kWideThreads, 0, s>>>(G, groups); break;
            case 1: o1_batch_profile<1><<<grid_batches, kWideThreads, 0, s>>>(G, groups); break;
            case 2: o1_batch_profile<2><<<grid_batches, kWideThreads, 0, s>>>(G, groups); break;
            case 3: o1_batch_profile<3><<<grid_batches, kWideThreads, 0, s>>>(G, groups); break;
            default: o1_batch_profile<4><<<grid_batches, kWideThreads, 0, s>>>(G, groups); break;
        }
    }
    if ((status = lh::check_launch("o1_batch_profile")) != LH_OK) return status;
    {
        long long deposit_waves = (long long)work * SL.segments * ((optics->bands + 1) / 2);
        const unsigned grid_deposit = (unsigned)((deposit_waves + kWideWaves - 1) / kWideWaves);
        lh::KernelMark mark("o1_batch_deposit", s);
        o1_batch_deposit<<<grid_deposit, kWideThreads, 0, s>>>(G, groups);
    }
    if ((status = lh::check_launch("o1_batch_deposit")) != LH_OK) return status;
    dim3 grid(work, SL.segments);
    {
        lh::KernelMark mark("o1_batches", s);
        switch (optics->max_power) {
            case 0: o1_batches<0><<<grid, kBatchThreads, 0, s>>>(G, groups); break;
            case 1: o1_batches<1><<<grid, kBatchThreads, 0, s>>>(G, groups); break;
            case 2: o1_batches<2><<<grid, kBatchThreads, 0, s>>>(G, groups); break;
            case 3: o1_batches<3><<<grid, kBatchThreads, 0, s>>>(G, groups); break;
            default: o1_batches<4><<<grid, kBatchThreads, 0, s>>>(G, groups); break;
        }
    }
    if ((status = lh::check_launch("o1_batches")) != LH_OK) return status;
    if (compute_bins) {
        dim3 sum_grid((unsigned)((optics->bands * optics->bins * 4 + kThreads - 1) / kThreads),
                      work);
        o1_sum_segments<<<sum_grid, kThreads, 0, s>>>(G);
        if ((status = lh::check_launch("o1_sum_segments")) != LH_OK) return status;
    }
    if (compute_bins) {
        long long waves = (long long)work * optics->bands;
        o1_finalize_bins<<<(unsigned)((waves + 3) / 4), kThreads, 0, s>>>(G);
        if ((status = lh::check_launch("o1_finalize_bins")) != LH_OK) return status;
    }
    o1_finalize<<<work, kThreads, 0, s>>>(G);
    return lh::check_launch("o1_finalize");
}

}  // extern "C"
