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

#include <cfloat>

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) segment_bounds(const double* receivers,
                                                           const double* starts,
                                                           const double* directions,
                                                           const double* lengths,
                                                           const double* start_ns, int n_el,
                                                           double flight, double* distance,
                                                           double* earliest) {
    __shared__ double best_d[kThreads];
    __shared__ double best_t[kThreads];
    int d = blockIdx.x;
    double rx = receivers[3 * d], ry = receivers[3 * d + 1], rz = receivers[3 * d + 2];
    double bd = DBL_MAX, bt = DBL_MAX;
    // four elements per thread at a time: every load is issued before the
    // arithmetic (minima do not depend on the order)
    constexpr int kBatch = 4;
    for (int i0 = threadIdx.x; i0 < n_el; i0 += kBatch * kThreads) {
        double sx[kBatch], sy[kBatch], sz[kBatch], ux[kBatch], uy[kBatch], uz[kBatch];
        double len[kBatch], t0[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            const int i = i0 + b * kThreads;
            const bool in = i < n_el;
            sx[b] = in ? starts[3 * i] : 0.0;
            sy[b] = in ? starts[3 * i + 1] : 0.0;
            sz[b] = in ? starts[3 * i + 2] : 0.0;
            ux[b] = in ? directions[3 * i] : 0.0;
            uy[b] = in ? directions[3 * i + 1] : 0.0;
            uz[b] = in ? directions[3 * i + 2] : 0.0;
            len[b] = in ? lengths[i] : 0.0;
            t0[b] = in && start_ns ? start_ns[i] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            if (i0 + b * kThreads >= n_el) break;
            double ox = rx - sx[b], oy = ry - sy[b], oz = rz - sz[b];
            double along = ox * ux[b] + oy * uy[b] + oz * uz[b];
            along = fmin(fmax(along, 0.0), len[b]);
            double dx = ox - along * ux[b], dy = oy - along * uy[b], dz = oz - along * uz[b];
            double value = dx * dx + dy * dy + dz * dz;
            bd = fmin(bd, value);
            if (start_ns) bt = fmin(bt, t0[b] + flight * sqrt(value));
        }
    }
    best_d[threadIdx.x] = bd;
    best_t[threadIdx.x] = bt;
    __syncthreads();
    for (int half = kThreads / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            best_d[threadIdx.x] = fmin(best_d[threadIdx.x], best_d[threadIdx.x + half]);
            best_t[threadIdx.x] = fmin(best_t[threadIdx.x], best_t[threadIdx.x + half]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (distance) distance[d] = sqrt(best_d[0]);
        if (earliest) earliest[d] = best_t[0];
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
    segment_bounds<<<n_receivers, kThreads, 0, s>>>(receivers, starts, directions, lengths,
                                                    start_ns, n_segments, flight, distance,
                                                    earliest);
    return lh::check_launch("segment_bounds");
}
