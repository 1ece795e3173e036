// Band fold of the order >= 2 pack onto the node bands (gfx950).
//
// Restates fold_bundle (band_fold.py:155-208): on every (subdivided) radial
// interval the folded kernel sum_lambda L[l, lambda, j] e^{-(mua_lambda -
// mua_ref) r} e^{i omega r (1/v_lambda - 1/v_ref)} S_lambda(t) is evaluated at
// four equally spaced nodes in log r and refitted as a cubic (the 4x4 inverse
// Vandermonde of each interval comes from the host). One thread per (interval,
// degree, pair, frequency) holds every band and node; FP64 arithmetic,
// complex64 output.
#include "lh_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBands = 64;
constexpr int kMaxRank = 16;

__global__ void __launch_bounds__(kThreads) fold_kernel(lh_fold_input in, float2* out) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int W = in.freq, P = in.pairs, D = in.degree + 1;
    int news = in.intervals * in.subdivide;
    long long total = (long long)news * D * P * W;
    if (g >= total) return;
    int w = (int)(g % W);
    long long rest = g / W;
    int pair = (int)(rest % P); rest /= P;
    int l = (int)(rest % D);
    int nw = (int)(rest / D);
    int old = nw / in.subdivide;
    double low = in.new_log_r[nw], high = in.new_log_r[nw + 1];
    const double nodes[4] = {0.0, 1.0 / 3, 2.0 / 3, 1.0};
    double at[4], radius[4], t_old[4];
    for (int n = 0; n < 4; ++n) {
        at[n] = low + nodes[n] * (high - low);
        t_old[n] = at[n] - in.log_r[old];
        radius[n] = exp(at[n]);
    }
    double fr[kMaxRank][4], fi[kMaxRank][4];
    for (int j = 0; j < in.rank; ++j)
        for (int n = 0; n < 4; ++n) fr[j][n] = fi[j][n] = 0.0;
    const float2* coef = static_cast<const float2*>(in.coefficients);
    double omega = in.omega[w];
    for (int m = 0; m < in.bands; ++m) {
        size_t base = ((((size_t)m * in.intervals + old) * D + l) * P + pair) * 4 * W + w;
        double cr[4], ci[4];
        for (int a = 0; a < 4; ++a) {
            float2 c = coef[base + (size_t)a * W];
            cr[a] = c.x;
            ci[a] = c.y;
        }
        double da = in.absorption[m] - in.absorption_ref;
        double dk = 1.0 / in.speed[m] - 1.0 / in.speed_ref;
        for (int n = 0; n < 4; ++n) {
            double t = t_old[n];
            double vr = ((cr[0] * t + cr[1]) * t + cr[2]) * t + cr[3];
            double vi = ((ci[0] * t + ci[1]) * t + ci[2]) * t + ci[3];
            double amp = exp(-da * radius[n]);
            double s, c;
            sincos(omega * radius[n] * dk, &s, &c);
            double xr = amp * c, xi = amp * s;
            double pr = vr * xr - vi * xi, pi = vr * xi + vi * xr;
            for (int j = 0; j < in.rank; ++j) {
                double mix = in.mixing[((size_t)l * in.bands + m) * in.rank + j];
                fr[j][n] += mix * pr;
                fi[j][n] += mix * pi;
            }
        }
    }
    const double* inv = in.inverse + (size_t)nw * 16;
    for (int j = 0; j < in.rank; ++j) {
        size_t obase = ((((size_t)j * news + nw) * D + l) * P + pair) * 4 * W + w;
        for (int a = 0; a < 4; ++a) {
            double sr = 0.0, si = 0.0;
            for (int n = 0; n < 4; ++n) {
                sr += inv[a * 4 + n] * fr[j][n];
                si += inv[a * 4 + n] * fi[j][n];
            }
            out[obase + (size_t)a * W] = make_float2((float)sr, (float)si);
        }
    }
}

}  // namespace

extern "C" int lh_fold_bundle(const lh_fold_input* input, void* out, void* stream) {
    LH_REQUIRE(input != nullptr && out != nullptr, "input and output must be set");
    const lh_fold_input& in = *input;
    LH_REQUIRE(in.coefficients && in.log_r && in.new_log_r && in.omega && in.absorption
                   && in.speed && in.mixing && in.inverse, "fold input pointers must be set");
    LH_REQUIRE(in.bands >= 1 && in.bands <= kMaxBands, "1..%d bands", kMaxBands);
    LH_REQUIRE(in.rank >= 1 && in.rank <= kMaxRank && in.rank <= in.bands, "1..%d node bands",
               kMaxRank);
    LH_REQUIRE(in.intervals >= 1 && in.subdivide >= 1 && in.degree >= 0 && in.pairs >= 1
                   && in.freq >= 1, "empty pack");
    LH_REQUIRE(in.speed_ref > 0, "reference speed must be positive");
    long long total = (long long)in.intervals * in.subdivide * (in.degree + 1) * in.pairs * in.freq;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fold_kernel<<<(unsigned)((total + kThreads - 1) / kThreads), kThreads, 0, s>>>(
        in, static_cast<float2*>(out));
    return lh::check_launch("fold_kernel");
}
