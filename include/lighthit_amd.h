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

#include <stddef.h>

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
 * Kernel timer (measurement only; no counterpart in the reference, which
 * times stages with perf_counter on the host). While enabled, the dominant
 * kernels of every entry point record a HIP event before and after their
 * launch on the stream they are launched on. lh_timing_read synchronises
 * those events and returns the summed milliseconds and the launch count of
 * one kernel name; lh_timing_reset releases every recorded event. Timing
 * adds two event records per launch; leave it off outside measurements.
 */
int lh_timing_enable(int on);
int lh_timing_read(const char* kernel, double* total_ms, long long* launches);
int lh_timing_reset(void);

/*
 * HIP graphs of a launch sequence (no counterpart in the reference: the
 * serving form replays the fused order-0 launch from one). lh_graph_begin
 * starts capturing `stream` (thread-local mode; not the null stream);
 * every launch the calling thread then enqueues on it is recorded, not run;
 * lh_graph_end stops the capture and instantiates it into *graph_exec;
 * lh_graph_launch replays it on a stream; lh_graph_destroy frees it. Only
 * sequences of kernel launches belong in a graph here (no memsets, no copies
 * from pageable memory).
 */
int lh_graph_begin(void* stream);
int lh_graph_end(void* stream, void** graph_exec);
int lh_graph_launch(void* graph_exec, void* stream);
int lh_graph_destroy(void* graph_exec);

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
 * Stage (c) in split precision (the precision form of the same call sites):
 * basis_hi, basis_lo (n_bins, 2*n_freq) FP16 with basis = (hi + lo) /
 * row_scale[row] (row_scale (n_bins) float, powers of two); the spectrum is
 * split into FP16 hi + lo per observation after a power-of-two scale, and the
 * products hi*hi + hi*lo + lo*hi accumulate in FP32 on FP16 MFMA. The phase
 * w * origin is formed and reduced in FP64. Other arguments as lh_readout_bins.
 */
int lh_readout_bins_split(const void* basis_hi, const void* basis_lo, const float* row_scale,
                          const double* omega, const void* spectrum, const double* origins,
                          int n_bins, int n_freq, int n_obs, float* bins, void* stream);

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
    int coefficient_format;     /* 0: complex64 in the layout above (FP32 VALU contraction);
                                   1: the precision form's MFMA layout: bfloat16 hi + lo,
                                   (bands, intervals, degree+1, freq_stride, KQ,
                                   [re hi, re lo, im hi, im lo], 4 powers), KQ = pairs
                                   rounded up to a multiple of 4 with zeros; contracted
                                   as hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16_bf16
                                   with FP32 accumulation */
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

/* Device workspace bytes lh_ge2_apply needs for a complex64 pack (0 when none). */
size_t lh_ge2_workspace_bytes(int bands, int freq, int receivers, int points);

/* Device workspace bytes of one lh_ge2_apply call with either coefficient format
 * (the MFMA form keeps its pair lists, records and per-pair spectra there). */
size_t lh_ge2_apply_workspace_bytes(const lh_ge2_pack* pack, const lh_ge2_source* source,
                                    int receivers, int freq);

/*
 * Descriptors are host structs holding device pointers. alpha (A+1) double
 * are the acceptance Legendre coefficients. The first `freq` frequencies are
 * evaluated (freq = 1: the zero-frequency charge). Points at a distance
 * outside [radial_low, radial_high] are skipped (radial_window_m).
 * out (freq, count) complex64. workspace: device memory of at least
 * lh_ge2_workspace_bytes (per-split partial sums; may be NULL when 0).
 */
int lh_ge2_apply(const lh_ge2_pack* pack, const lh_ge2_source* source,
                 const lh_ge2_receivers* receivers, const double* alpha, int freq,
                 double radial_low, double radial_high, void* out, void* workspace,
                 size_t workspace_bytes, void* stream);

/*
 * Stage (a): order-0 (direct) light to the first glass entry of every OM.
 *
 * Replaces PreparedBallistic(kernel, source, config).charge()
 * (_surface_ballistic.py:20-98, reached from transport_prompt's order-0
 * block, prompt.py:637-655): the candidate (OM, element) pairs of
 * interval_candidates (_ballistic_interval_numba.py:450-519) and the
 * first-entry integral of each, summed per OM. Spherical receivers; regular
 * pairs only (element line farther than the glass radius from the OM centre,
 * no constant-phase panel); other pairs set flag bit 1 and are skipped.
 */
typedef struct lh_order0_optics {
    const double* panel_edges;       /* (panels+1) nm, from prepare_optics */
    const double* phase_nodes;       /* (panels+1) */
    const double* group_nodes;       /* (panels+1) */
    const double* extinction_nodes;  /* (panels+1) 1/m */
    const double* efficiency;        /* (panels, efficiency_columns) cubic in q */
    const int* major;                /* (major_count) panel-edge indices of the medium
                                        nodes inside the band, both band ends included */
    const double* alpha;             /* (acceptance_count) acceptance Legendre coefficients */
    int panels, efficiency_columns, major_count, acceptance_count;
    double phase_min, phase_max;     /* extrema of phase_nodes */
} lh_order0_optics;

typedef struct lh_order0_oms {
    const double* positions;   /* (count, 3) m */
    const double* looks;       /* (count, 3) -orientation */
    const double* areas;       /* (count) effective area m^2 */
    const double* radii;       /* (count) glass radius m */
    int count;
} lh_order0_oms;

typedef struct lh_order0_elements {
    const double* start_m;        /* (count, 3) */
    const double* direction;      /* (count, 3) unit */
    const double* length_m;       /* (count) */
    const double* coefficient0;   /* (count) C0 */
    const double* coefficient2;   /* (count) C2 = -C0 / beta^2 */
    const double* beta;           /* (count) */
    const double* start_ns;       /* (count) */
    const double* end_ns;         /* (count) */
    int count;
} lh_order0_elements;

/* Bytes of device workspace lh_order0_charge needs (flags + per-chunk partial sums). */
size_t lh_order0_workspace_bytes(int n_oms, int n_elements);

/*
 * charge (oms->count) double, written. workspace: device memory of at least
 * lh_order0_workspace_bytes; its first int holds flags after the call:
 * bit 0 an emitting segment intersects the glass (the reference raises
 * ValueError, _surface_ballistic.py:52-53), bit 1 a pair outside the
 * supported geometry was skipped.
 */
int lh_order0_charge(const lh_order0_optics* optics, const lh_order0_oms* oms,
                     const lh_order0_elements* elements, double* charge, void* workspace,
                     size_t workspace_bytes, void* stream);

/*
 * Fused order-0 charge of several events on one set of OMs (the serving
 * form): one persistent launch over the (event, OM, chunk) items of all of
 * them, in place of one lh_order0_charge per event. The reference evaluates
 * one event per call (PreparedBallistic.charge, _surface_ballistic.py:91-112,
 * reached once per ShowerTransport call, shower/api.py:176-190); every
 * event's charges equal its own lh_order0_charge result bit for bit.
 *
 * Two steps, so the second can be captured in a HIP graph once per capacity
 * and replayed: lh_order0_batch_set writes the table of this batch's events
 * (at most max_events, their element chunks at most max_chunks; outputs
 * charges[k] (oms->count) double each) into the workspace with a kernel;
 * lh_order0_charge_batch runs the launch chain, every grid sized by
 * (oms->count, max_events, max_chunks) and every event size read from the
 * table on the device. lh_order0_batch_chunks gives the chunks of a batch
 * (from its element counts), lh_order0_batch_workspace_bytes the workspace of
 * a capacity. The flag words of the workspace report as for lh_order0_charge,
 * for the batch as a whole.
 */
size_t lh_order0_batch_workspace_bytes(int n_oms, int max_events, long long max_chunks);
long long lh_order0_batch_chunks(const int* element_counts, int n_events);
int lh_order0_batch_set(const lh_order0_oms* oms, const lh_order0_elements* events,
                        double* const* charges, int n_events, int max_events, long long max_chunks,
                        void* workspace, size_t workspace_bytes, void* stream);
int lh_order0_charge_batch(const lh_order0_optics* optics, const lh_order0_oms* oms,
                           int max_events, long long max_chunks, void* workspace,
                           size_t workspace_bytes, void* stream);

/*
 * lh_order0_charge without a workspace argument and without any device
 * allocation. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The workspace lives in a fixed device array of the
 * library, 64 MiB per GPU (lh_order0_charge_noalloc_max_bytes), held from the
 * first call on a GPU until the process ends; a call whose
 * lh_order0_workspace_bytes(oms->count, elements->count) exceeds it fails with
 * LH_ERR_ARGUMENT. flags (device int, may be NULL) receives the flag word that
 * lh_order0_charge leaves in its workspace. Every call on one GPU shares that
 * array: calls that overlap on one GPU (concurrent streams or host threads)
 * race and must be serialised by the caller. The result equals
 * lh_order0_charge's bit for bit (same kernels).
 */
size_t lh_order0_charge_noalloc_max_bytes(void);
int lh_order0_charge_noalloc(const lh_order0_optics* optics, const lh_order0_oms* oms,
                             const lh_order0_elements* elements, double* charge, int* flags,
                             void* stream);

/*
 * Order-0 arrival bins of the OMs in `oms` (normally the active subset).
 *
 * Replaces PreparedBallistic.bins(ids, origins, edges) (_surface_ballistic.py:
 * 100-112, called from prompt.py:863-868): arrival time
 * t = start_ns + tau s + d_entry n_g / c - origin, bins half-open [e_k, e_k+1),
 * light outside [e_0, e_n) dropped. Each azimuth chord is split exactly at the
 * bin edges it crosses.
 * origins (oms->count) double; bin_edges (n_bins+1) double, increasing;
 * bins (oms->count, n_bins) double, written; n_bins <= 2048.
 */
size_t lh_order0_bins_workspace_bytes(int n_oms, int n_elements, int n_bins);
int lh_order0_bins(const lh_order0_optics* optics, const lh_order0_oms* oms,
                   const lh_order0_elements* elements, const double* origins,
                   const double* bin_edges, int n_bins, double* bins, void* workspace,
                   size_t workspace_bytes, void* stream);

/*
 * Diagnostics for the order-0 cost model; no counterpart in the reference.
 * counter (one device unsigned 64-bit word, or NULL to stop counting): every
 * later order-0 launch of the process adds the (theta, azimuth) quadrature
 * nodes it evaluates. A process-wide setting, not for concurrent callers.
 */
int lh_order0_count_nodes(unsigned long long* counter);

/*
 * Measurement only (no counterpart): the quadrature rules of every later
 * order-0 launch of the process. s_nodes 0 (the default) chooses per pair:
 * for the charge 3 Gauss nodes per s piece when the element is at most a
 * tenth of the glass radius long and 6 otherwise, for the bins 6; 2, 3, 4 or
 * 6 force that rule. azimuth_nodes 0 (the default) chooses per pair: 4
 * azimuth nodes (per azimuth piece of the bins) when the glass radius is at
 * most 1/20 of the OM's distance from the element's line and 8 otherwise; 4
 * or 8 force that rule. A caller compares rule orders with it (the quadrature's
 * error estimate). A process-wide setting, not for concurrent callers.
 */
int lh_order0_set_rules(int s_nodes, int azimuth_nodes);

/*
 * Order >= 2 source: the axial-slice compile of one light field.
 *
 * Replaces compile_axial_slices (_shower_basis.py:188-645) as the shower route
 * calls it per node band (shower/ge2.py:49-55): chord quadrature nodes, cubic
 * (or linear) Lagrange weights on axis slices, a linear (or nearest) time
 * lattice, a lattice row per occupied (slice, tick), thin/wide slice decision,
 * deposit of weight * moment * P_l(cone) Y_lm(u) per row, the phase sum
 * e^{i omega t} per slice and the five-point transverse stencil. Two calls:
 * lh_compile_layout builds the lattice and the point layout and writes
 * info = {slices, wide slices, points, rows, error flags} (device ints);
 * after reading info the caller sizes the deposit buffer (rows * (moments *
 * channels + 2 * n_freq) doubles: the rows' deposits, then their phase table)
 * and the channels, and calls lh_compile_channels with the
 * same workspace. Element arrays are shared by every band; a band is
 * selected by its photon scales and phase index (photons = scale0 * C0 -
 * scale2 * (-C2), cone = 1 / (beta n), elements with beta n <= 1 or mask 0
 * skipped), or photons/cone_cosine are given directly.
 */
typedef struct lh_compile_field {
    const double* start_m;       /* (count, 3) */
    const double* direction;     /* (count, 3) unit */
    const double* length_m;      /* (count) */
    const double* start_ns;      /* (count) */
    const double* end_ns;        /* (count) */
    const double* coefficient0;  /* (count) or NULL when photons is given */
    const double* coefficient2;  /* (count) or NULL when photons is given */
    const double* beta;          /* (count) or NULL when photons is given */
    const double* photons;       /* (count) direct photon weights, or NULL */
    const double* cone_cosine;   /* (count) direct cone cosines, or NULL */
    const unsigned char* mask;   /* (count) elements taken, or NULL for all */
    double scale0, scale2, phase_index;
    int count;
} lh_compile_field;

typedef struct lh_compile_lattice {
    double centre[3], axis[3], first[3], second[3];   /* shower frame */
    double z_origin_m, reference_ns, cell_m, time_step_ns, stencil_m, stencil_threshold_m;
    int degree, azimuthal_degree;
    int axial_order;      /* 1 or 3 */
    int moments;          /* 1 (collapse) or 5 (stencil) */
    int element_order;    /* chord Gauss nodes, 1 or 2 */
    int time_deposit;     /* 0 nearest, 1 linear, 2 charge only (one tick) */
    int kmax;             /* ticks per slice in the row key */
    int slice0;           /* lowest slice index of the lattice */
    int positions;        /* slice positions of the lattice */
} lh_compile_lattice;

size_t lh_compile_workspace_bytes(const lh_compile_lattice* lattice, int count);
int lh_compile_layout(const lh_compile_field* field, const lh_compile_lattice* lattice,
                      int* info, double* points, void* workspace, size_t workspace_bytes,
                      void* stream);
int lh_compile_channels(const lh_compile_field* field, const lh_compile_lattice* lattice,
                        const double* omega, int n_freq, double scale, int rows,
                        double* deposit, void* channels, void* workspace,
                        size_t workspace_bytes, void* stream);

/*
 * Order 1 (single scattering): grouped pencils.
 *
 * Replaces the grouped branch of _order1 (prompt.py:1020-1258):
 * compress (prompt.py:1317-1402 with compress_subsegments and finish_pencils,
 * _prompt_numba.py:1094-1230, 2247-2354), compress_screen + pack_pencils
 * (_prompt_screen.py:44-93), grouped_pencils_order1
 * (_prompt_numba.py:1750-2065) and screen_order1 (_prompt_screen.py:170-477).
 * Compression is two calls: lh_order1_compress_keys writes the number of
 * pencil keys to info[0]; lh_order1_compress_pencils writes that many pencils
 * (screen grids, parent_face > 0: parents and children in preorder, at most
 * twice as many, count in info[1]). The grouped pass is two calls:
 * lh_order1_group_keys writes the number of (tau, r, theta) groups to info[0]
 * and the bits of its compact per-OM keys to info[1] (it returns after the
 * stream has drained); lh_order1_group_charge takes both, a scratch buffer of
 * lh_order1_group_scratch_bytes(pencils, oms, optics, groups, pool_nodes) and
 * writes charge (and bins) of the listed OMs. The scratch beyond the fixed
 * part is a pool of path nodes (24 bytes each; pool_nodes -1 asks for the
 * default size, 0 for none): when every batch's path fits it, the batches run
 * as wide kernels (one wave per batch, then one per (OM, segment, wavelength));
 * otherwise one workgroup per (OM, segment) does it all. Both give the same
 * sums in the same order. Its optional work_counts (4 device
 * counters, added to) receive the number of profile terms (path nodes x
 * groups), time panels, bin events and group members ((pencil, OM) entries
 * in groups) it evaluated. Each pair shares its
 * workspace. lh_order1_set_tables uploads the 24-term Taylor table of
 * I_k(z) (3 x 24) and the 48-node inverse-angle rule (tangent, weight sec^2).
 */
typedef struct lh_order1_chords {
    const double* start_m;       /* (elements, 3) */
    const double* direction;     /* (elements, 3) unit */
    const double* length_m;
    const double* coefficient0;
    const double* coefficient2;
    const double* cone_cosine;   /* (elements) cone of the unit */
    const double* start_ns;
    const double* end_ns;
    const int* index;            /* (count) element of every chord of the unit, or NULL */
    const long long* sub_offset; /* (count + 1) first sub-segment of every chord */
    int count;
} lh_order1_chords;

typedef struct lh_order1_grid {
    double origin[3];
    long long dims[3];
    double cell_m;
    int pixel_face, ring_nodes, parent_face;   /* parent_face 0: no screen hierarchy */
    long long subsegments;
} lh_order1_grid;

typedef struct lh_order1_pencils {
    double* position;    /* (count, 3) */
    double* direction;   /* (count, 3) */
    double* w0;
    double* w2;
    double* time;
    double* major;       /* (count, 3) */
    double* var_major;
    double* var_minor;
    double* pos_cov;     /* (count, 6) xx yy zz xy xz yz */
    long long* parent;   /* screen: -2 for parents, else the parent's index; or NULL */
    long long* next;     /* screen: next pencil after a parent's subtree; or NULL */
    int count;
} lh_order1_pencils;

typedef struct lh_order1_oms {
    const double* positions;   /* (N, 3) */
    const double* looks;       /* (N, 3) */
    const double* areas;       /* (N) */
    const double* origins;     /* (N) time origins */
    const int* work;           /* (work_count) OMs evaluated */
    int work_count;
} lh_order1_oms;

typedef struct lh_order1_optics {
    const double* mus;
    const double* mut;
    const double* speed;
    const double* s0w;
    const double* s2w;           /* (bands) */
    const double* edges;         /* (bins + 1) relative time edges */
    const double* eedge;         /* (bands, bins + 1) e^{kappa edge} */
    const double* itab;          /* (bands, bins, 3) I_k(kappa width) */
    const double* coef;          /* (n_mono) acceptance monomials */
    const int* ipow;
    const int* jpow;
    int bands, bins, n_mono, max_power, direct_bins, max_path_nodes;
    double g, y_g, path_step, path_step_tail, tail_exponent, smoothing, mut_min, speed_max;
    /* equal bins: 1 / width (0 when the edges are not equally spaced); bin
     * lookups then start at floor((t - edges[0]) / width) */
    double bin_inverse_width;
} lh_order1_optics;

typedef struct lh_order1_merge {
    double theta_lo, theta_ratio, tau_step, log_r_step, speed_ref;
} lh_order1_merge;

int lh_order1_set_tables(const double* taylor, const double* tangent,
                         const double* weighted_sec2);
size_t lh_order1_compress_workspace_bytes(const lh_order1_chords* chords,
                                          const lh_order1_grid* grid);
int lh_order1_compress_keys(const lh_order1_chords* chords, const lh_order1_grid* grid, int* info,
                            void* workspace, size_t workspace_bytes, void* stream);
int lh_order1_compress_pencils(const lh_order1_chords* chords, const lh_order1_grid* grid,
                               int keys, lh_order1_pencils* out, int* info, void* workspace,
                               size_t workspace_bytes, void* stream);
size_t lh_order1_grouped_workspace_bytes(const lh_order1_pencils* pencils,
                                         const lh_order1_oms* oms,
                                         const lh_order1_optics* optics);
int lh_order1_group_keys(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                         const lh_order1_optics* optics, const lh_order1_merge* merge, int* info,
                         void* workspace, size_t workspace_bytes, void* stream);
size_t lh_order1_group_scratch_bytes(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                                     const lh_order1_optics* optics, int groups,
                                     long long pool_nodes);
int lh_order1_group_charge(const lh_order1_pencils* pencils, const lh_order1_oms* oms,
                           const lh_order1_optics* optics, const lh_order1_merge* merge,
                           int groups, int key_bits, void* scratch, size_t scratch_bytes,
                           int compute_bins, double* charge, double* bins,
                           unsigned long long* work_counts, void* workspace,
                           size_t workspace_bytes, void* stream);

/*
 * Near-field clamp of the order >= 2 light (near_field.py:91-173 with
 * _cone_many_looks, muon/track_ge2_primitive.py:612-660): for every element
 * portion closer to an OM than the tables' radial minimum, the spectrum
 * (relative to the OM's time origin) of its clamped order >= 2 light,
 *   out[q, w] = density e^{i w phase_ns} sum_nodes w_i e^{-i w u_i clock}
 *               sum_bands scale_b(R) e^{i w r / v_b} sum_l h_bl sum_pairs
 *               rot_l(mu) alpha_lambda Y_lambda,mu(look') cubic_b,j,l,pair(t),
 * r = hypot(impact, u_i), R = max(r, radial_low) for the amplitude
 * (attenuation, 1/(4 pi R^2), spline), r for the flight phase; nodes with
 * r < clamp_low or r > radial_high are left out. pack is the per-wavelength
 * (unfolded) bundle; portion q's frame is (ex toward the OM, ey = axis x ex,
 * axis = element direction); look is the OM look in that frame. out is
 * (count, nfreq) complex128 (interleaved doubles).
 */
typedef struct lh_near_portions {
    const double* impact;      /* (count) OM distance from the element line [m] */
    const double* look;        /* (count, 3) OM look in the portion frame */
    const double* weight;      /* (count, bands) wavelength weight x efficiency x area */
    const double* harmonics;   /* (count, bands, degree + 1) cone harmonics per unit C0 */
    const double* phase_ns;    /* (count) closest-point emission time minus the OM origin */
    const double* clock;       /* (count) element time per metre [ns/m] */
    const double* density;     /* (count) C0 per metre */
    const int* node_start;     /* (count + 1) first quadrature node of each portion */
    const double* u;           /* (nodes) OM minus source along the axis [m] */
    const double* w;           /* (nodes) quadrature weights [m] */
    int count;
} lh_near_portions;

int lh_near_field(const lh_ge2_pack* pack, const lh_near_portions* portions, const double* alpha,
                  int nfreq, double clamp_low, double radial_low, double radial_high, double* out,
                  void* stream);

/*
 * Band fold of the order >= 2 pack (one-time table preparation).
 *
 * Replaces fold_bundle (band_fold.py:155-208): coefficients (bands,
 * intervals, degree+1, pairs, 4, freq) complex64 of the per-wavelength pack
 * -> out (rank, intervals*subdivide, degree+1, pairs, 4, freq) complex64 with
 * common attenuation absorption_ref and flight speed speed_ref. mixing is
 * BandFold.interpolation (degree+1, bands, rank); inverse holds the 4x4
 * inverse Vandermonde matrix of every new interval (node values -> t^3..t^0).
 */
typedef struct lh_fold_input {
    const void* coefficients;
    const double* log_r;        /* (intervals + 1) */
    const double* new_log_r;    /* (intervals * subdivide + 1) */
    const double* omega;        /* (freq) */
    const double* absorption;   /* (bands) */
    const double* speed;        /* (bands) */
    const double* mixing;       /* (degree + 1, bands, rank) */
    const double* inverse;      /* (intervals * subdivide, 4, 4) */
    double absorption_ref, speed_ref;
    int bands, intervals, degree, pairs, freq, rank, subdivide;
} lh_fold_input;

int lh_fold_bundle(const lh_fold_input* input, void* out, void* stream);

/*
 * Per-OM bounds of a source's segments: distance (n_receivers) to the nearest
 * point of any segment (min_distance_to_segments, _prompt_numba.py:2221) and,
 * when earliest is set, the earliest arrival start_ns + flight * distance
 * (earliest_element_arrival, _shower_basis.py:51-78; flight = n_g,min / c).
 * Either output may be NULL.
 */
int lh_segment_bounds(const double* receivers, int n_receivers, const double* starts,
                      const double* directions, const double* lengths, const double* start_ns,
                      int n_segments, double flight, double* distance, double* earliest,
                      void* stream);

/*
 * Window report of one compiled band (shower_basis.py:408-439
 * _radial_window_report): for every OM, the basis points closer than low
 * (near), farther than high (far) and inside [low, high] (used), and the
 * estimate of the light beyond high,
 *   tail = scale * area * sum_far weight |ch_p| att(r) / (4 pi r^2),
 * with att(r) = exp(interp(r, radius, log_factor)) (np.interp, clamped).
 * points are frame coordinates (first, second, axis); frame holds the
 * centre, axis, first and second vectors (12 doubles); channels is the
 * band's complex64 channel array, read at point p's first element
 * (channel_stride complex values apart). The caller passes
 * weight = sqrt(4 pi) / efficiency and scale = 10 peak_acceptance efficiency.
 */
typedef struct lh_window_input {
    const double* points;
    const void* channels;
    int count, channel_stride;
    const double* frame;
    const double* positions;
    const double* areas;
    int n_oms;
    double low, high, weight, scale;
    const double* radius;
    const double* log_factor;
    int table;
} lh_window_input;

int lh_window_report(const lh_window_input* input, int* used, int* far, int* near, double* tail,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTHIT_AMD_H */
