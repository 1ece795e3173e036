// Per-OM segment bounds of a source (gfx950): the nearest distance to any
// emitting segment (min_distance_to_segments, _prompt_numba.py:2221-2243) and
// the earliest possible arrival start_ns + n_g,min |x - segment| / c
// (earliest_element_arrival, _shower_basis.py:51-78). One workgroup per OM;
// threads stride over segments and reduce the minimum (exact, order-free).
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
    for (int i = threadIdx.x; i < n_el; i += kThreads) {
        double ox = rx - starts[3 * i], oy = ry - starts[3 * i + 1], oz = rz - starts[3 * i + 2];
        double ux = directions[3 * i], uy = directions[3 * i + 1], uz = directions[3 * i + 2];
        double along = ox * ux + oy * uy + oz * uz;
        along = fmin(fmax(along, 0.0), lengths[i]);
        double dx = ox - along * ux, dy = oy - along * uy, dz = oz - along * uz;
        double value = dx * dx + dy * dy + dz * dz;
        bd = fmin(bd, value);
        if (start_ns) bt = fmin(bt, start_ns[i] + flight * sqrt(value));
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

}  // namespace

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
