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

#ifdef __cplusplus
}
#endif

#endif /* LIGHTHIT_AMD_H */
