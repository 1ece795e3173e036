// Per-OM geometry of the shower route (gfx950), one workgroup per OM:
//  * segment bounds: the nearest distance to any emitting segment
//    (min_distance_to_segments, _prompt_numba.py:2221-2243) and the earliest
//    possible arrival start_ns + n_g,min |x - segment| / c
//    (earliest_element_arrival, _shower_basis.py:51-78); threads stride over
//    segments and reduce the minimum (exact, order-free);
//  * window report of a compiled band (shower_basis.py:408-439): basis points
//    inside, below and beyond the radial window of the order >=2 tables and
//    the far-tail estimate of the light beyond it; threads stride over points
//    and reduce in a fixed tree order, so the tail does not depend on
//    scheduling.
#include "lh_common.h"

#include <algorithm>
#include <cfloat>

namespace {

constexpr int kThreads = 256;

constexpr int kBoundOms = 8;          // OMs per workgroup of segment_bounds
constexpr int kBoundThreads = 256;
constexpr int kBoundBlocks = 2048;    // target workgroups of one launch
constexpr int kBoundMinElements = 1024;   // least elements of one split

// Orderable 64-bit encoding of a double (unsigned order = numeric order).
__device__ inline unsigned long long order_bits(double x) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

__device__ inline double from_order_bits(unsigned long long b) {
    return __longlong_as_double((long long)((b >> 63) ? (b & 0x7fffffffffffffffULL) : ~b));
}

// The outputs hold the running minima as order_bits encodings until
// segment_bounds_finish decodes them; all ones lies above every encoding.
__global__ void __launch_bounds__(kThreads) segment_bounds_init(int n_receivers,
                                                                double* distance,
                                                                double* earliest) {
    const int d = blockIdx.x * kThreads + threadIdx.x;
    if (d >= n_receivers) return;
    if (distance) reinterpret_cast<unsigned long long*>(distance)[d] = ~0ULL;
    if (earliest) reinterpret_cast<unsigned long long*>(earliest)[d] = ~0ULL;
}

// Nearest approach and earliest arrival of kBoundOms OMs over one split of
// the elements (blockIdx.y of gridDim.y equal splits): each element a
// workgroup loads serves all its OMs, and the splits keep the chip full when
// the OMs are few. Minima do not depend on the order, so the workgroup minima
// go to the outputs with atomic minima on their ordered encodings. An element
// whose start time is not below the running earliest arrival cannot lower it
// (the flight term is >= 0), so its square root is skipped.
__global__ void __launch_bounds__(kBoundThreads) segment_bounds(const double* receivers,
                                                                int n_receivers,
                                                                const double* starts,
                                                                const double* directions,
                                                                const double* lengths,
                                                                const double* start_ns, int n_el,
                                                                double flight, double* distance,
                                                                double* earliest) {
    __shared__ double part[kBoundThreads / 64][2 * kBoundOms];
    const int first = blockIdx.x * kBoundOms;
    const int per = (n_el + gridDim.y - 1) / gridDim.y;
    const int lo = blockIdx.y * per, hi = min(n_el, lo + per);
    double rx[kBoundOms], ry[kBoundOms], rz[kBoundOms], bd[kBoundOms], bt[kBoundOms];
#pragma unroll
    for (int k = 0; k < kBoundOms; ++k) {
        const int d = min(first + k, n_receivers - 1);
        rx[k] = receivers[3 * d];
        ry[k] = receivers[3 * d + 1];
        rz[k] = receivers[3 * d + 2];
        bd[k] = DBL_MAX;
        bt[k] = DBL_MAX;
    }
    for (int i = lo + threadIdx.x; i < hi; i += kBoundThreads) {
        const double sx = starts[3 * i], sy = starts[3 * i + 1], sz = starts[3 * i + 2];
        const double ux = directions[3 * i], uy = directions[3 * i + 1], uz = directions[3 * i + 2];
        const double len = lengths[i];
        const double t0 = start_ns ? start_ns[i] : 0.0;
#pragma unroll
        for (int k = 0; k < kBoundOms; ++k) {
            const double ox = rx[k] - sx, oy = ry[k] - sy, oz = rz[k] - sz;
            double along = ox * ux + oy * uy + oz * uz;
            along = fmin(fmax(along, 0.0), len);
            const double dx = ox - along * ux, dy = oy - along * uy, dz = oz - along * uz;
            const double value = dx * dx + dy * dy + dz * dz;
            bd[k] = fmin(bd[k], value);
            if (start_ns && t0 < bt[k]) bt[k] = fmin(bt[k], t0 + flight * sqrt(value));
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kBoundOms; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            bd[k] = fmin(bd[k], __shfl_xor(bd[k], off, 64));
            bt[k] = fmin(bt[k], __shfl_xor(bt[k], off, 64));
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kBoundOms; ++k) {
            part[wave][k] = bd[k];
            part[wave][kBoundOms + k] = bt[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kBoundOms) {
        double best = part[0][threadIdx.x];
        for (int w = 1; w < kBoundThreads / 64; ++w) best = fmin(best, part[w][threadIdx.x]);
        const int k = threadIdx.x % kBoundOms;
        const int d = first + k;
        if (d < n_receivers && best < DBL_MAX) {
            double* out = threadIdx.x < kBoundOms ? distance : earliest;
            if (out) atomicMin(reinterpret_cast<unsigned long long*>(out) + d, order_bits(best));
        }
    }
}

// Decodes the minima: the distance is the square root of the nearest squared
// distance.
__global__ void __launch_bounds__(kThreads) segment_bounds_finish(int n_receivers,
                                                                  double* distance,
                                                                  double* earliest) {
    const int d = blockIdx.x * kThreads + threadIdx.x;
    if (d >= n_receivers) return;
    if (distance) {
        const unsigned long long b = reinterpret_cast<const unsigned long long*>(distance)[d];
        distance[d] = sqrt(from_order_bits(b));
    }
    if (earliest) {
        const unsigned long long b = reinterpret_cast<const unsigned long long*>(earliest)[d];
        earliest[d] = from_order_bits(b);
    }
}

// np.interp(x, xp, fp) on an increasing table: clamped at both ends.
__device__ double interp(double x, const double* xp, const double* fp, int n) {
    if (x <= xp[0]) return fp[0];
    if (x >= xp[n - 1]) return fp[n - 1];
    int low = 0, high = n - 1;
    while (high - low > 1) {
        int middle = (low + high) >> 1;
        if (xp[middle] <= x) low = middle; else high = middle;
    }
    double slope = (fp[low + 1] - fp[low]) / (xp[low + 1] - xp[low]);
    return slope * (x - xp[low]) + fp[low];
}

__global__ void __launch_bounds__(kThreads) window_report(lh_window_input in, int* used, int* far,
                                                          int* near, double* tail) {
    __shared__ double s_sum[kThreads];
    __shared__ int s_near[kThreads];
    __shared__ int s_far[kThreads];
    const int d = blockIdx.x;
    const double px = in.positions[3 * d], py = in.positions[3 * d + 1],
                 pz = in.positions[3 * d + 2];
    const double* f = in.frame;
    const float2* ch = static_cast<const float2*>(in.channels);
    double sum = 0.0;
    int n_near = 0, n_far = 0;
    const double four_pi = 12.566370614359172953850573533118;
    for (int i = threadIdx.x; i < in.count; i += kThreads) {
        const double a = in.points[3 * i], b = in.points[3 * i + 1], c = in.points[3 * i + 2];
        // detector coordinates: centre + z axis + t1 first + t2 second
        const double x = f[0] + c * f[3] + a * f[6] + b * f[9];
        const double y = f[1] + c * f[4] + a * f[7] + b * f[10];
        const double z = f[2] + c * f[5] + a * f[8] + b * f[11];
        const double dx = x - px, dy = y - py, dz = z - pz;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r < in.low) {
            ++n_near;
        } else if (r > in.high) {
            ++n_far;
            const float2 v = ch[(size_t)i * in.channel_stride];
            const double photons = in.weight * (double)hypotf(v.x, v.y);
            const double att = exp(interp(r, in.radius, in.log_factor, in.table));
            sum += photons * att / (four_pi * r * r);
        }
    }
    s_sum[threadIdx.x] = sum;
    s_near[threadIdx.x] = n_near;
    s_far[threadIdx.x] = n_far;
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + half];
            s_near[threadIdx.x] += s_near[threadIdx.x + half];
            s_far[threadIdx.x] += s_far[threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        near[d] = s_near[0];
        far[d] = s_far[0];
        used[d] = in.count - s_near[0] - s_far[0];
        tail[d] = in.scale * in.areas[d] * s_sum[0];
    }
}

}  // namespace

extern "C" int lh_window_report(const lh_window_input* input, int* used, int* far, int* near,
                                double* tail, void* stream) {
    LH_REQUIRE(input != nullptr, "lh_window_report: null input");
    const lh_window_input& in = *input;
    LH_REQUIRE(in.n_oms >= 0 && in.count >= 0 && in.table >= 2 && in.channel_stride >= 1,
               "lh_window_report: need n_oms >= 0, count >= 0, table >= 2, channel_stride >= 1");
    if (in.n_oms == 0) return LH_OK;
    LH_REQUIRE(in.positions && in.areas && in.radius && in.log_factor && used && far && near &&
                   tail && (in.count == 0 || (in.points && in.channels)),
               "lh_window_report: null pointer");
    LH_REQUIRE(in.low <= in.high, "lh_window_report: low > high");
    hipStream_t s = static_cast<hipStream_t>(stream);
    window_report<<<in.n_oms, kThreads, 0, s>>>(in, used, far, near, tail);
    return lh::check_launch("window_report");
}

extern "C" int lh_segment_bounds(const double* receivers, int n_receivers, const double* starts,
                                 const double* directions, const double* lengths,
                                 const double* start_ns, int n_segments, double flight,
                                 double* distance, double* earliest, void* stream) {
    LH_REQUIRE(receivers && starts && directions && lengths, "geometry pointers must be set");
    LH_REQUIRE(n_receivers >= 1 && n_segments >= 1, "need receivers and segments");
    LH_REQUIRE(earliest == nullptr || start_ns != nullptr, "earliest needs start_ns");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!distance && !earliest) return LH_OK;
    const int groups = (n_receivers + kBoundOms - 1) / kBoundOms;
    const int splits = std::max(1, std::min((kBoundBlocks + groups - 1) / groups,
                                            n_segments / kBoundMinElements));
    const unsigned flat = (unsigned)((n_receivers + kThreads - 1) / kThreads);
    segment_bounds_init<<<flat, kThreads, 0, s>>>(n_receivers, distance, earliest);
    int status = lh::check_launch("segment_bounds_init");
    if (status != LH_OK) return status;
    segment_bounds<<<dim3(groups, splits), kBoundThreads, 0, s>>>(
        receivers, n_receivers, starts, directions, lengths, start_ns, n_segments, flight,
        distance, earliest);
    if ((status = lh::check_launch("segment_bounds")) != LH_OK) return status;
    segment_bounds_finish<<<flat, kThreads, 0, s>>>(n_receivers, distance, earliest);
    return lh::check_launch("segment_bounds_finish");
}
