/*
 * lighthit_amd — C-ABI of the MI355X (gfx950) kernels behind
 * ShowerTransport(kernel, data)(shower).
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer owned by the caller, unless
 *     its comment says "host". No entry point allocates device memory.
 *   - Complex arrays are interleaved (re, im) pairs: complex64 = 2 x float.
 *   - Array shapes are given row-major, (slowest, ..., fastest).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Work is only enqueued: the caller synchronises.
 *   - Return value: LH_OK, or an error code whose message is available from
 *     lh_last_error() (per calling thread).
 *       LH_ERR_ARGUMENT  maps to the reference's ValueError
 *       LH_ERR_HIP       maps to RuntimeError (a HIP runtime failure)
 */
#ifndef LIGHTHIT_AMD_H
#define LIGHTHIT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define LH_ABI_VERSION 1

enum lh_status {
    LH_OK = 0,
    LH_ERR_ARGUMENT = 1,
    LH_ERR_HIP = 2
};

/* Message of the last failed call on this thread ("" if none). */
const char* lh_last_error(void);
void lh_clear_error(void);
int lh_abi_version(void);

/*
 * Stage (c): order >=2 time bins of OMs from their spectra.
 *
 * Replaces, for the order >=2 column of the shower readout,
 *   relative = spectrum * exp(-i w origin)              shower/api.py:461
 *   CubicSpline(w, relative)(readout_w)  (re and im)    shower/api.py:462-465
 *   inverse_bins(readout_w, relative, edges)            shower/api.py:466-467,
 *                                                       readout.py:18-40
 * basis    (n_bins, 2*n_freq) float: [Re M | -Im M], M = K S, K the exact
 *          per-bin integral of exp(-i w t) with trapezoid weights on the
 *          readout grid and S the not-a-knot spline matrix (readout grid x
 *          table grid); built once per grid by the host.
 * omega    (n_freq) double, table frequencies [rad/ns].
 * spectrum (n_freq, n_obs) complex64, absolute-time spectra.
 * origins  (n_obs) double time origins [ns]; NULL = spectra are already
 *          relative (plain inverse_bins, readout.py:18).
 * bins     (n_obs, n_bins) float, written.
 */
int lh_readout_bins(const float* basis, const double* omega, const void* spectrum,
                    const double* origins, int n_bins, int n_freq, int n_obs,
                    float* bins, void* stream);

/*
 * Stage (b): order >=2 response of OMs to a fused multi-band axial source,
 * applied with the band-folded kernel pack.
 *
 * Replaces apply_multi_spectral(bundle, sources, receivers_m, look_directions,
 * alpha, out_scale, radial_window_m=...) (_shower_basis.py:1126-1198) with its
 * compiled core _apply_multi_spectral (_shower_basis.py:895-1050), as called by
 * apply_folded (shower/ge2.py:99-119) for the predicted, exact zero-frequency
 * and full-frequency >=2 charges of the shower route (shower/api.py:347-352,
 * 405-406, 434-435). Every band's channels keep their own emission-time
 * reference (the reference rounds band-aligned channels instead); the result
 * is the absolute-time spectrum (frequency, OM).
 *
 * Pairs of the pack are the canonical (detector degree, |mu|) list of the
 * reference's CouplingTable (directional.py:161-169): lambda = 0..A, mu =
 * 0..lambda. Built for acceptance degree A in {2, 3} and azimuthal degree
 * min(source azimuthal degree, pack degree) in 0..4.
 */
typedef struct lh_ge2_pack {
    const void* coefficients;   /* (bands, intervals, degree+1, pairs, 4, freq_stride)
                                   complex64, cubic powers t^3..t^0 in log r */
    const double* log_r;        /* (intervals+1) */
    const double* absorption;   /* (bands) common attenuation [1/m] */
    const double* speed;        /* (bands) flight speed [m/ns] */
    const double* omega;        /* (freq_stride) [rad/ns] */
    int bands, intervals, degree, acceptance_degree, freq_stride;
    int flight;                 /* 1: radial splines carry the flight phase */
} lh_ge2_pack;

typedef struct lh_ge2_source {
    const double* points_m;      /* (points, 3) in the shower frame (first, second, axis) */
    const void* channels_data;   /* (bands, points, channels, freq_stride) complex64 */
    const double* reference_ns;  /* (bands) emission-time reference of each band */
    int points, channels, degree, azimuthal_degree, freq_stride;
} lh_ge2_source;

typedef struct lh_ge2_receivers {
    const double* positions;     /* (count, 3) in the shower frame */
    const double* looks;         /* (count, 3) unit look directions in the shower frame */
    const double* out_scale;     /* (bands, count) detector scale (OM area) */
    int count;
} lh_ge2_receivers;

/*
 * Descriptors are host structs holding device pointers. alpha (A+1) double
 * are the acceptance Legendre coefficients. The first `freq` frequencies are
 * evaluated (freq = 1: the zero-frequency charge). Points at a distance
 * outside [radial_low, radial_high] are skipped (radial_window_m).
 * out (freq, count) complex64.
 */
int lh_ge2_apply(const lh_ge2_pack* pack, const lh_ge2_source* source,
                 const lh_ge2_receivers* receivers, const double* alpha, int freq,
                 double radial_low, double radial_high, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTHIT_AMD_H */
