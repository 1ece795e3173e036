"""Benchmark of the ShowerTransport call on one MI355X.

The headline is the serving + mixed-precision form: events of one workload
served by worker processes on one GPU, each calling
``ShowerTransport(kernel, data)(shower)`` with ``precision="mixed"`` (the
multi-frequency order >=2 apply as bfloat16 hi + lo products and the readout
as FP16 hi + lo products on MFMA, FP32 accumulation; orders 0 and 1, every
charge and every selection decision in FP32/FP64). One step = one batch of
``--processes`` events (``events_per_step``); ``--steps`` batches are timed
after ``--warmup`` batches. Beside it: the plain FP32 form and the plain mixed
form, one call at a time; the charge error of the mixed form per order against
the FP64 oracle route (the committed detector-scale fixture); a serving stream
of events with uneven element counts at Poisson arrival rates; GPU time per
stage and kernel (HIP events on the launch stream); the rooflines of the three
stages (order 0 on the vector ALUs, the order >=2 apply and the readout on
MFMA) with counter figures from rocprofv3 passes; and the FP64 oracle on the
host cores, one thread and all threads, over a bounded sample.

Workload (default): the detector-scale event: 2304 OMs (8 clusters x 8
strings x 36 OMs), a 50,000-element Gaussian shower of seed 11 with a 1500 m
charged track, 8 wavelength nodes, 1 ns bins over -60..745 ns, 65 RTE and 321
readout frequencies, degree 8, azimuthal degree 2, radial window 2..150 m,
order-1 selection radius 300 m. The order >=2 pack is built from the RTE
tables (lighthit_amd.tables, cached on disk between processes; ``--pack
synthetic`` takes the seeded made-up pack instead).
Sources are resident in HBM before the timed region. ``--gpus`` above 1
exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
FP16_PEAK_TFLOPS = 2500.0
#: SIMDs of the MI355X (256 CUs x 4): the denominator of MFMA busy fractions.
SIMDS = 1024
#: XCDs of the MI355X: GRBM_GUI_ACTIVE reads as the sum of their active cycles.
XCDS = 8

# FP32 operations of one (theta, azimuth) node of the order-0 integrand at
# acceptance degree 2 (each add, multiply, divide, square root, exponential
# or sine counted once): per azimuth node two 6-term sine series (28), their
# arguments and the cap gap (20), the entry distance (11), cos/sin of the
# azimuth (9), the two acceptance arguments and polynomials (31), the
# exponential weight and the sums (8); plus the theta-node set-up (two
# sincos, the index and wavelength of the cone, the efficiency cubic, the
# spectral and Jacobian factors, the cap half-width: ~100) shared by its 8
# azimuth nodes.
ORDER0_NODE_FLOPS = 107 + 100 // 8
O1_MEMBER_BYTES = 22 * 8 + 4

# FP64 operations per unit of the order-1 batch kernel (each add, multiply,
# divide, square root or exponential counted once):
#   profile term (one path node x one group): scattering geometry 16, the two
#     cubic acceptance polynomials 28, Henyey-Greenstein weight 8, sums 4;
#   time panel: quadratic coefficients 30, two attenuation factors 4, the
#     exponential-quadratic integral 64 (three 9-term Taylor series, the
#     length FP64 resolves at the panel widths here, and the combination);
#   bin event: one more exponential-quadratic integral over a bin piece 70.
O1_PROFILE_TERM_FLOPS = 56
O1_PANEL_FLOPS = 98
O1_EVENT_FLOPS = 70
# FP64 operations of one group member of the order-1 group sums: the pencil's
# frame at the OM (distance, direction cosines, angle: ~40), its key (log,
# floor, divisions: ~8), the shape factor or radial rule (~30), the weights
# and moments at polynomial degree 1 (~18) and its one add per value into its
# group's sum (14).
O1_MEMBER_FLOPS = 110

#: Kernels the library times with HIP events on their stream.
TIMED_KERNELS = ("order0_pairs<false>", "order0_pairs<true>", "o1_batch_profile",
                 "o1_batch_deposit", "o1_batches", "o1_group_sums", "ge2_wide", "ge2_narrow",
                 "ge2_mfma_apply", "readout_kernel", "readout_split_kernel", "compile_deposit",
                 "compile_phase")

#: The three stages the rooflines report: kernel, engine.
STAGE_KERNELS = {"order0": "order0_pairs<false>", "ge2_apply": "ge2_mfma_apply",
                 "readout": "readout_split_kernel"}


def build_route(name, precision="fp32", pack="rte"):
    """Workload, route (tables resident) in ``precision`` and the route kernel;
    ``pack`` "rte" (the order >=2 pack built from the RTE tables, cached on
    disk) or "synthetic"."""
    from dataclasses import replace

    from lighthit_amd.route import ShowerTransport
    from lighthit_amd.synthetic import workload_route
    workload, kernel, route = workload_route(name, pack=pack)
    if precision != route.config.precision:
        route = ShowerTransport(route.kernel, route.data,
                                config=replace(route.config, precision=precision))
    return workload, kernel, route


def route_factory(name, precision="fp32", pack="rte"):
    """The route alone, built the same way in a serving worker process."""
    return build_route(name, precision, pack)[2]


# ----------------------------------------------------------- roofline models

def order0_bytes(n_elements, n_oms, bins):
    """Algorithmic HBM bytes of one order-0 launch: every element read once
    (12 FP32 words: start, direction, length, times, two coefficients, beta,
    weight), every OM once (position, look, area, radius: 9 words) and the
    FP64 result written once (a charge, or ``bins`` per OM)."""
    return 48 * n_elements + 36 * n_oms + 8 * n_oms * max(1, bins)


def ge2_flops_per_term(degree, azimuthal, acceptance, pairs):
    """FP32 flops of one (OM, source point, band, frequency) term of the full
    order >=2 apply: the cubic radial interpolation of every (multipole,
    pair) coefficient (4 complex-by-real multiply-adds, 4 flops each), the
    pair sum into the detector-side vector (4 flops per pair), the Wigner
    rotation of the source channels (complex-by-real, 4 flops per matrix
    element), their complex dot product (8 flops per term) and the phase and
    scale (16 flops)."""
    cubic = (degree + 1) * pairs * 4 * 4
    pair_sum = (degree + 1) * pairs * 4
    rotation = sum((2 * min(l, acceptance) + 1) * (2 * min(l, azimuthal) + 1)
                   for l in range(degree + 1)) * 4
    dot = sum(2 * min(l, acceptance) + 1 for l in range(degree + 1)) * 8
    return cubic + pair_sum + rotation + dot + 16


def ge2_window_terms(route, positions, window):
    """(OM, source point, band) triples inside the radial window for the OMs
    ``positions`` and the node-band points of the route's last compile."""
    sources = route._compiled["sources"][0]
    low, high = window
    total = 0
    for item in sources:
        points = item.to_points_m()
        distance = np.linalg.norm(positions[:, None, :] - points[None, :, :], axis=2)
        total += int(np.count_nonzero((distance >= low) & (distance <= high)))
    return total


def _cpu_sample(workload, kernel, route, response, stride):
    """One pass of the FP64 oracle over the bounded sample of the event; seconds
    per stage (order 0 scaled from every ``stride``-th element to all)."""
    from oracle import ge2 as oracle_ge2
    from oracle import host as oracle_host
    from oracle import order0 as oracle_order0
    from oracle import order1 as oracle_order1
    from oracle import readout as oracle_readout
    from oracle import route as oracle_route

    d = workload.detector
    elements = workload.elements
    tables = route.tables()
    fold = tables["fold"]
    folded = dict(tables["folded"])
    folded["coefficients"] = np.asarray(folded["coefficients"].cpu().numpy(), np.complex64)
    zero = dict(folded, coefficients=np.ascontiguousarray(folded["coefficients"][..., :1]),
                omega=np.asarray(folded["omega"][:1]))
    alpha = np.asarray(kernel.acceptance()["alpha"], float)
    optics = oracle_host.prepare_optics(kernel.medium, d, kernel.config.wavelength_range_nm)
    radii = d.receiver_surface.radii(len(d))
    looks = -np.asarray(d.orientations, float)
    level1 = np.asarray(response.metadata["prompt_order1_level"])
    level2 = np.asarray(response.metadata["ge2_charge_level"])
    work1 = np.flatnonzero(level1 > 0)
    candidates = np.flatnonzero(level2 == 2)
    active = np.flatnonzero(response.active)
    full = np.flatnonzero(response.active & (level2 > 0))
    origins = np.asarray(response.time_origin_ns, float)
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    omega = np.asarray(kernel.config.omega_per_ns, float)
    sample = elements.subset(np.arange(len(elements)) % stride == 0)
    seconds = {}
    began = time.perf_counter()
    oracle_order0.first_entry(d.positions_m[work1], looks[work1], d.effective_area_m2[work1],
                              radii[work1], sample, optics, alpha, rel_tol=2e-6)
    if len(active):
        oracle_order0.first_entry(d.positions_m[active], looks[active],
                                  d.effective_area_m2[active], radii[active], sample, optics,
                                  alpha, origins=origins[active], edges=edges, rel_tol=2e-6)
    seconds["order0"] = (time.perf_counter() - began) * len(elements) / len(sample)
    began = time.perf_counter()
    spectral = oracle_order1.spectral_tables(kernel.medium, d, kernel.wavelength)
    charge = response.charge_components_pe
    oracle_order1.spectral_order1(elements, d.positions_m, looks, d.effective_area_m2,
                                  origins, work1, spectral, alpha, float(kernel.medium.g),
                                  edges, float(kernel.config.threshold_pe), charge[:, 0],
                                  2.0 * np.abs(charge[:, 2]), oracle_order1.PromptSettings())
    seconds["order1"] = time.perf_counter() - began
    began = time.perf_counter()
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    max_phase = float(np.max(kernel.medium.sample(wavelengths)["phase_index"]))
    frame = oracle_host.axis_frame(elements.field(0, phase_index=max_phase))
    sources = oracle_route._node_sources(
        elements, wavelengths, fold, omega, frame, int(kernel.config.source_degree),
        int(kernel.config.azimuthal_degree), route.config.basis)
    seconds["compile"] = time.perf_counter() - began
    began = time.perf_counter()
    window = tuple(kernel.config.radial_range_m)
    pos, area = d.positions_m, d.effective_area_m2
    zero_sources = [copy.copy(s) for s in sources]
    for s in zero_sources:
        object.__setattr__(s, "channels", np.ascontiguousarray(s.channels[:, :, :1]))
    if len(candidates):
        oracle_ge2.apply_folded(zero, zero_sources, pos[candidates], looks[candidates],
                                alpha, area[candidates], window)
    spectrum = np.zeros((len(omega), len(full)), complex)
    if len(full):
        spectrum = oracle_ge2.apply_folded(folded, sources, pos[full], looks[full], alpha,
                                           area[full], window)
    seconds["order2"] = time.perf_counter() - began
    began = time.perf_counter()
    if len(full):
        oracle_readout.route_readout(omega, spectrum, origins[full],
                                     kernel.config.readout_omega_per_ns, edges)
    seconds["readout"] = time.perf_counter() - began
    return seconds, len(sample), (len(work1), len(active), len(candidates), len(full))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, kernel, route, response, fraction, repeats):
    """The FP64 oracle on the host cores over a bounded sample of the same event,
    on one thread and on every thread of this process's CPU share, the median of
    ``repeats`` passes each: order 0 (charge at the selected OMs, bins at the
    active OMs) on every ``1/fraction``-th element scaled to all elements;
    order 1, the node-band compile, the exact zero-frequency and full order >=2
    apply and the readout in full on the OM sets the GPU route chose."""
    from threadpoolctl import threadpool_limits
    stride = max(1, int(round(1 / fraction)))
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads_all = int(os.environ.get("OMP_NUM_THREADS", 0) or share)
    runs = {}
    for threads in (1, threads_all):
        passes = []
        with threadpool_limits(limits=threads):
            for _ in range(repeats):
                seconds, n_sample, sets = _cpu_sample(workload, kernel, route, response, stride)
                passes.append(seconds)
        totals = [sum(p.values()) for p in passes]
        median = passes[int(np.argsort(totals)[len(totals) // 2])]
        runs[threads] = {"events_per_s": 1.0 / float(np.median(totals)),
                         "seconds_per_event_median": float(np.median(totals)),
                         "seconds_per_event_all": totals, "stage_seconds": median}
    work1, active, candidates, full = sets
    n = len(workload.elements)
    law_s = 8.0 + 0.18e-3 * n
    best = runs[threads_all]
    return {"value": best["events_per_s"], "unit": "events/s", "cores": threads_all,
            "kind": "port",
            "sample": (f"FP64 oracle (C with OpenMP and numpy), one event, median of {repeats} "
                       f"passes: order 0 on every {stride}th element ({n_sample} of {n}) at the "
                       f"reference tolerance 2e-6, scaled to all elements, at {work1} selected "
                       f"and {active} active OMs; order 1 on {work1} OMs, node-band compile, "
                       f"order >=2 zero frequency at {candidates} and full apply and readout at "
                       f"{full} OMs in full; windows and predictor not timed"),
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "cpu_share": share,
            "threads": {str(t): r for t, r in runs.items()},
            "note": ("an unoptimised restatement: the oracle keeps every candidate (OM, "
                     "element) pair of order 0 at the reference's tolerance, with no bounded "
                     "omission; the reference publishes ~8 s + 0.18 ms per step on one CPU "
                     f"thread (README.md:516-518), {law_s:.1f} s for this event's {n} steps "
                     f"({1.0 / law_s:.3f} events/s), for a 1440-OM array")}


# ------------------------------------------------------------ serving stream

def _compact(response):
    rows = np.flatnonzero(response.components_pe.reshape(len(response.components_pe), -1)
                          .any(axis=1))
    return response.charge_components_pe.copy(), rows, response.components_pe[rows].copy()


def plain_stream(route, stream):
    """The stream's events one plain call at a time (host arrays, upload inside):
    per-event service times and the compact results the served events must equal."""
    import torch
    import warnings
    results, times = [], []
    began = time.perf_counter()
    for elements in stream:
        start = time.perf_counter()
        try:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)
                response = route(elements)
            results.append(_compact(response))
        except ValueError as error:
            results.append(("error", str(error)))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - start)
    elapsed = time.perf_counter() - began
    lat = np.asarray(times) * 1e3
    return {"results": results,
            "report": {"events_per_s": len(stream) / elapsed,
                       "latency_ms": {"p50": float(np.percentile(lat, 50)),
                                      "p99": float(np.percentile(lat, 99))},
                       "errors": sum(1 for r in results if isinstance(r[0], str))}}


def served_stream(server, stream, plain, rates):
    """The stream's events submitted at Poisson arrival times to ``server``, at
    each rate: throughput, latency from arrival to result, and whether every
    result equals the plain call's."""
    from lighthit_amd.serve import latency_summary
    from lighthit_amd.synthetic import poisson_arrivals
    report = {"events": len(stream), "plain": plain["report"],
              "element_counts": {"min": min(len(e) for e in stream),
                                 "median": int(np.median([len(e) for e in stream])),
                                 "max": max(len(e) for e in stream)},
              "rates": []}
    for rate in rates:
        print(f"[bench] serving stream at {rate:g} events/s", file=sys.stderr, flush=True)
        times = poisson_arrivals(rate, len(stream))
        futures = []
        start_at = time.perf_counter()
        for elements, at in zip(stream, times):
            delay = start_at + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futures.append(server.submit(elements))
        same, errors = True, 0
        for f, want in zip(futures, plain["results"]):
            want_error = isinstance(want[0], str)
            try:
                got = _compact(f.result())
            except ValueError:
                errors += 1
                same &= want_error
                continue
            same &= (not want_error) and all(np.array_equal(a, b) for a, b in zip(got, want))
        finished = max(f.submitted + f.latency_s for f in futures)
        report["rates"].append({"offered_events_per_s": rate,
                                "events_per_s": len(futures) / (finished - start_at),
                                "latency_ms": latency_summary(futures),
                                "errors": errors, "same_as_plain": bool(same)})
    return report


# ------------------------------------------------------------------- driver

def ge2_pairs(route, positions, window):
    """(OM, source point) pairs inside the radial window for the OMs
    ``positions``: every band of the node-band compile shares the points."""
    sources = route._compiled["sources"][0]
    points = sources[0].to_points_m()
    distance = np.linalg.norm(positions[:, None, :] - points[None, :, :], axis=2)
    return int(np.count_nonzero((distance >= window[0]) & (distance <= window[1])))


def window_work(route, workload, response):
    """Order >=2 work of the event the route compiled last, at the OMs of the full
    apply: (OM, point, band) terms and (OM, point) pairs inside the radial window."""
    d = workload.detector
    level2 = np.asarray(response.metadata["ge2_charge_level"])
    full = np.flatnonzero(response.active & (level2 > 0))
    return {"terms": ge2_window_terms(route, d.positions_m[full], workload.radial_range_m),
            "pairs": ge2_pairs(route, d.positions_m[full], workload.radial_range_m),
            "bands": len(route._compiled["sources"][0])}


def roofline_of(name, kernel_ms, workload, response, route, counts=None, launches=1,
                nodes=None, window=None):
    """Roofline entry of kernel ``name`` at its average launch duration ``kernel_ms``:
    the algorithmic work of one launch (a model stated in ``algorithmic``) over
    that duration, against the peak of the engine the kernel runs on."""
    d = workload.detector
    n_el = len(workload.elements)
    seconds = kernel_ms * 1e-3
    if name == "order0_pairs<false>" and nodes is not None:
        flops = nodes * ORDER0_NODE_FLOPS
        achieved = flops / seconds / 1e12
        hbm = order0_bytes(n_el, len(d), 0) / seconds / 1e9
        return {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "algorithmic": f"per launch {nodes:.4g} (theta, azimuth) quadrature nodes "
                               f"(counted in the kernel) x {ORDER0_NODE_FLOPS} FP32 flops",
                "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": hbm / HBM_PEAK_GBS,
                        "algorithmic": f"48 B/element x {n_el} + 36 B/OM x {len(d)} + 8 B/OM"},
                "note": "FP32 vector work (sines, exponentials, divisions per node); priced "
                        "against the FP32 vector peak"}
    if name.startswith("order0_pairs"):
        bins = len(workload.edges_ns) - 1 if "true" in name else 0
        oms = len(d) if bins == 0 else int(np.count_nonzero(response.active))
        achieved = order0_bytes(n_el, oms, bins) / seconds / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "algorithmic": f"48 B/element x {n_el} + 36 B/OM x {oms} + 8 B x {max(1, bins)} "
                               f"outputs/OM"}
    if name in ("o1_batch_deposit", "o1_batch_profile", "o1_batches") and counts is not None:
        terms, panels, events = (c / launches for c in counts[:3])
        if name == "o1_batch_profile":
            flops = O1_PROFILE_TERM_FLOPS * terms
            algorithmic = (f"per launch {terms:.4g} profile terms x {O1_PROFILE_TERM_FLOPS} "
                           f"FP64 flops (counted in the kernel)")
        elif name == "o1_batch_deposit":
            flops = O1_PANEL_FLOPS * panels + O1_EVENT_FLOPS * events
            algorithmic = (f"per launch {panels:.4g} panels x {O1_PANEL_FLOPS} + {events:.4g} "
                           f"bin events x {O1_EVENT_FLOPS} FP64 flops (counted in the kernel)")
        else:
            flops = (O1_PROFILE_TERM_FLOPS * terms + O1_PANEL_FLOPS * panels
                     + O1_EVENT_FLOPS * events)
            algorithmic = (f"per launch {terms:.4g} profile terms x {O1_PROFILE_TERM_FLOPS} "
                           f"+ {panels:.4g} panels x {O1_PANEL_FLOPS} + {events:.4g} bin "
                           f"events x {O1_EVENT_FLOPS} FP64 flops (counted in the kernel)")
        achieved = flops / seconds / 1e12
        return {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "algorithmic": algorithmic,
                "note": "FP64 vector work (no matrix form); priced against the FP64 peak"}
    if name == "o1_group_sums" and counts is not None and len(counts) > 3:
        members = counts[3] / launches
        achieved = members * O1_MEMBER_FLOPS / seconds / 1e12
        return {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "algorithmic": f"per launch {members:.4g} group members (counted in the kernel) "
                               f"x {O1_MEMBER_FLOPS} FP64 flops",
                "note": "dependent FP64 geometry per member after a gather of its pencil "
                        "(latency-bound); priced against the FP64 vector peak"}
    if name in ("ge2_wide", "ge2_mfma_apply") and window is not None:
        from lighthit_amd.ge2 import canonical_pairs
        pack = route.tables()["device_pack"]
        bands = window["bands"]
        nfreq = len(workload.omega_per_ns)
        if name == "ge2_wide":
            terms = window["terms"]
            per = ge2_flops_per_term(workload.source_degree, workload.azimuthal_degree,
                                     pack.acceptance_degree,
                                     len(canonical_pairs(pack.acceptance_degree)[0]))
            flops = float(terms) * nfreq * per
            achieved = flops / seconds / 1e12
            return {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                    "algorithmic": f"{terms} (OM, point, band) terms in the window x "
                                   f"{nfreq} frequencies x {per} flops"}
        pairs = window["pairs"]
        npair = len(canonical_pairs(pack.acceptance_degree)[0])
        m = min(workload.azimuthal_degree, workload.source_degree)
        rows = sum(2 * min(l, m) + 1 for l in range(workload.source_degree + 1))
        # H = Q C per (pair, degree): rows x (pairs x 4 powers) x (re, im), 3 split products
        flops = float(pairs) * bands * nfreq * rows * npair * 4 * 2 * 2 * 3
        achieved = flops / seconds / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                "algorithmic": f"{pairs} (OM, point) pairs in the window x {bands} bands x "
                               f"{nfreq} frequencies x {rows} detector rows x {npair * 4} "
                               "(multipole pair, power) terms x 2 (re, im) x 2 flops x 3 "
                               "bfloat16 products (hi hi, hi lo, lo hi)",
                "note": "the channel epilogue (complex FP32 products per row) is not priced"}
    if name in ("readout_split_kernel", "readout_kernel"):
        level2 = np.asarray(response.metadata["ge2_charge_level"])
        full = int(np.count_nonzero(response.active & (level2 > 0)))
        bins = len(workload.edges_ns) - 1
        nfreq = len(workload.omega_per_ns)
        products = 3 if name == "readout_split_kernel" else 1
        flops = 2.0 * bins * (2 * nfreq) * full * products
        peak = FP16_PEAK_TFLOPS if products == 3 else FP32_PEAK_TFLOPS
        achieved = flops / seconds / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak,
                "algorithmic": f"{bins} bins x {2 * nfreq} (re, im) table frequencies x {full} "
                               f"OMs x 2 flops x {products} "
                               + ("FP16 products (hi hi, hi lo, lo hi)" if products == 3
                                  else "FP32 product")}
    return None


# ------------------------------------------------------------ counter passes

def _read_counters(folder):
    """{kernel name: [{counter: value} per dispatch]} from a rocprofv3 CSV folder:
    values summed over the counter's dimensions (GRBM counters: their maximum)."""
    import csv
    import glob
    per = {}
    names = {}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                key = (path, row["Dispatch_Id"])
                names[key] = row["Kernel_Name"]
                counter = row["Counter_Name"]
                value = float(row["Counter_Value"])
                entry = per.setdefault(key, {})
                if counter.startswith("GRBM_"):
                    entry[counter] = max(entry.get(counter, 0.0), value)
                else:
                    entry[counter] = entry.get(counter, 0.0) + value
    out = {}
    for key, entry in per.items():
        out.setdefault(names[key], []).append(entry)
    return out


def _counter_mean(table, name_part, counter):
    values = [e[counter] for k, rows in table.items() if name_part in k
              for e in rows if counter in e]
    return sum(values) / len(values) if values else None


def counter_passes(workload_flag, precision, dominant, keep_dir=None, pack="rte"):
    """rocprofv3 counter passes over one event of the plain route (a child
    process each): HBM bytes per launch of ``dominant`` (FETCH_SIZE and
    WRITE_SIZE, each scaled by what it reads for a plain 256 MiB copy) and
    the MFMA busy fraction, vector-ALU busy fraction and executed flops of
    the stage kernels."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    base = os.path.abspath(keep_dir) if keep_dir else tempfile.mkdtemp(prefix="lh_counters_")
    os.makedirs(base, exist_ok=True)
    me = os.path.abspath(__file__)
    passes = {
        "fetch": ("FETCH_SIZE", True),
        "write": ("WRITE_SIZE", True),
        "engines": ("SQ_VALU_MFMA_BUSY_CYCLES SQ_ACTIVE_INST_VALU SQ_INSTS_VALU_MFMA_MOPS_BF16 "
                    "SQ_INSTS_VALU_MFMA_MOPS_F16 GRBM_GUI_ACTIVE", False),
    }
    tables = {}
    for tag, (counters, calibrate) in passes.items():
        runs = [("bench", ["--profile-run", "--workload", workload_flag, "--precision", precision,
                           "--pack", pack])]
        if calibrate:
            runs.append(("copy", ["--copy-kernel-mib", "256"]))
        for kind, args in runs:
            out = os.path.join(base, f"{tag}_{kind}")
            cmd = ["rocprofv3", "--pmc", *counters.split(), "--output-format", "csv",
                   "-d", out, "-o", "run", "--", sys.executable, me, *args]
            print(f"[bench] counter pass {tag} {kind}", file=sys.stderr, flush=True)
            done = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd="/tmp",
                                  env=dict(os.environ, TMPDIR="/tmp"))
            if done.returncode != 0:
                return None, f"rocprofv3 {tag} {kind} failed ({done.returncode})"
            tables[(tag, kind)] = _read_counters(out)
    result = {}
    for tag, counter in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        copy_kib = _counter_mean(tables[(tag, "copy")], "copyBuffer", counter) or \
            _counter_mean(tables[(tag, "copy")], "elementwise", counter)
        kernel_kib = _counter_mean(tables[(tag, "bench")], dominant.split("<")[0], counter)
        if not copy_kib or kernel_kib is None:
            return None, f"no {counter} samples"
        scale = 256 * 1024 / copy_kib    # the copy moves 256 MiB each way
        result[counter] = (kernel_kib * scale * 1024, scale, kernel_kib)
    traffic = result["FETCH_SIZE"][0] + result["WRITE_SIZE"][0]
    note = (f"{dominant}: FETCH_SIZE {result['FETCH_SIZE'][2]:.0f} KiB x"
            f"{result['FETCH_SIZE'][1]:.3f} + WRITE_SIZE {result['WRITE_SIZE'][2]:.0f} KiB x"
            f"{result['WRITE_SIZE'][1]:.3f} per dispatch (copy-calibrated), bytes per launch")
    engines = {}
    table = tables[("engines", "bench")]
    for stage, kernel_name in dict(STAGE_KERNELS, dominant=dominant).items():
        part = kernel_name.replace("<false>", "<false").split("<")[0]
        rows = [e for k, v in table.items()
                if part in k and ("<true" not in k or "true" in kernel_name) for e in v]
        if not rows:
            continue
        mean = {c: sum(e.get(c, 0.0) for e in rows) / len(rows) for c in rows[0]}
        active = (mean.get("GRBM_GUI_ACTIVE") or 0.0) / XCDS
        # executed MFMA flops (MOPS x 512, padding included); busy fractions as
        # rocprof's MfmaUtil and VALUBusy define them, per GPU clock of the dispatch
        entry = {"dispatches": len(rows),
                 "mfma_flops_executed": 512 * (mean.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0)
                                               + mean.get("SQ_INSTS_VALU_MFMA_MOPS_F16", 0.0))}
        if active:
            entry["mfma_busy"] = mean.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (active * SIMDS)
            entry["valu_busy"] = mean.get("SQ_ACTIVE_INST_VALU", 0.0) / (active * SIMDS // 4)
        engines[stage] = {"kernel": kernel_name, **entry}
    return {"traffic": traffic, "traffic_note": note, "engines": engines}, None


def time_plain(route, elements, steps, warmup, *, timers=True):
    """``steps`` events one call at a time (sources resident, a fresh object per
    step, so the per-event compile always runs): throughput, latency, GPU time
    per stage and kernel, and whether every call returned the same result."""
    import torch
    from lighthit_amd._lib import KernelTimer, StageClock
    from lighthit_amd.route import ResidentShower
    first = None
    for _ in range(warmup):
        first = route(ResidentShower(elements))
    resident = [ResidentShower(elements) for _ in range(steps)]
    torch.cuda.synchronize()
    latencies, same = [], True
    with StageClock() as clock, KernelTimer() as timer:
        torch.cuda.synchronize()
        began = time.perf_counter()
        for source in resident:
            start = time.perf_counter()
            response = route(source)
            torch.cuda.synchronize()
            latencies.append(time.perf_counter() - start)
            if first is not None:
                same &= bool(np.array_equal(response.charge_components_pe,
                                            first.charge_components_pe)
                             and np.array_equal(response.components_pe, first.components_pe))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - began
        stages = {k: v / steps for k, v in clock.milliseconds().items()}
        kernels = {}
        for name in TIMED_KERNELS:
            total, launches = timer.read(name)
            if launches:
                kernels[name] = {"ms_per_launch": total / launches,
                                 "launches_per_event": launches / steps,
                                 "ms_per_event": total / steps}
        timer.reset()
    lat = np.asarray(latencies) * 1e3
    return {"events_per_s": steps / elapsed, "ms_per_event": 1e3 * elapsed / steps,
            "latency_ms": {"p50": float(np.percentile(lat, 50)),
                           "p99": float(np.percentile(lat, 99)), "max": float(lat.max())},
            "repeatable": bool(same), "stage_gpu_ms": stages, "kernel_gpu_ms": kernels,
            "response": response}


def precision_report(workload, response, pack="rte"):
    """Charge error of ``response`` against the FP64 oracle route on the same
    event and pack (the committed detector-scale fixture) per order: max and
    median relative error over the OMs whose FP64 charge of that order is above
    1e-6 of the largest, the bins' L1 error over the active OMs, and whether
    the active set is the same."""
    path = os.path.join(ROOT, "tests", "golden", "detector_scale_route.npz" if pack == "synthetic"
                        else f"detector_scale_route_{pack}.npz")
    if workload.name.split(":")[0] != "plain-detector-scale" or not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        want = {k: f[k] for k in ("charge", "active", "active_ids", "bins")}
    report = {"same_active": bool(np.array_equal(response.active, want["active"]))}
    ids = want["active_ids"]
    for order, name in ((0, "0"), (1, "1"), (2, ">=2")):
        ref = want["charge"][:, order]
        got = response.charge_components_pe[:, order]
        lit = np.abs(ref) > 1e-6 * np.abs(ref).max()
        rel = np.abs(got[lit] - ref[lit]) / np.abs(ref[lit])
        b_ref = want["bins"][:, :, order].astype(float)
        b_got = response.components_pe[ids, :, order]
        l1 = float(np.abs(b_got - b_ref).sum() / max(np.abs(b_ref).sum(), 1e-300))
        report[name] = {"charge_max_rel": float(rel.max()),
                        "charge_median_rel": float(np.median(rel)),
                        "bins_l1_rel": l1, "oms": int(lit.sum())}
    return report


def serve_processes(workload_flag, precision, detector, elements, processes, steps, warmup,
                    stream=None, stream_plain=None, rates=(), pack="rte"):
    """``steps`` batches of ``processes`` events served by ``processes`` worker
    processes (sources staged in device memory before the timed region), then
    Poisson arrivals at 70% of that throughput, then the serving stream."""
    import functools

    from lighthit_amd.serve import ProcessShowerServer, latency_summary
    events = steps * processes
    with ProcessShowerServer(functools.partial(route_factory, workload_flag, precision, pack),
                             processes=processes, detector=detector,
                             device="cuda:0") as server:
        server.map([elements for _ in range(max(1, warmup) * processes)])
        staged = [server.stage(elements) for _ in range(events)]
        began = time.perf_counter()
        futures = [server.submit(handle) for handle in staged]
        results = [f.result() for f in futures]
        elapsed = time.perf_counter() - began
        for handle in staged:
            server.release(handle)
        rate = 0.7 * events / elapsed
        staged = [server.stage(elements) for _ in range(events)]
        gaps = np.random.default_rng(23).exponential(1.0 / rate, len(staged))
        arrivals = []
        start_at = time.perf_counter()
        for handle, at in zip(staged, np.cumsum(gaps)):
            delay = start_at + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            arrivals.append(server.submit(handle))
        for f in arrivals:
            f.result()
        for handle in staged:
            server.release(handle)
        stream_report = None
        if stream is not None:
            stream_report = served_stream(server, stream, stream_plain, rates)
        stats = dict(server.stats)
    return {"events_per_s": events / elapsed, "events": events, "steps": steps,
            "events_per_step": processes, "processes": processes,
            "ms_per_step": 1e3 * elapsed / steps, "ms_per_event": 1e3 * elapsed / events,
            "latency_ms_all_submitted_at_once": latency_summary(futures),
            "latency_ms_poisson": dict(latency_summary(arrivals), rate_events_per_s=rate),
            "out_of_memory_retries": int(stats.get("out_of_memory_retries", 0)),
            "largest_peak_bytes": int(stats.get("largest_peak_bytes", 0)),
            "results": results, "stream": stream_report}


def main():
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=60,
                        help="timed steps: batches of --processes served events")
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", choices=["detector", "small", "heavy"], default="detector")
    parser.add_argument("--pack", choices=["rte", "synthetic"], default="rte",
                        help="order >=2 pack: built from the RTE tables (cached on disk) or the "
                             "seeded synthetic one")
    parser.add_argument("--precision", choices=["mixed", "fp32"], default="mixed",
                        help="precision of the served form (the headline)")
    parser.add_argument("--processes", type=int, default=8,
                        help="worker processes of the served form")
    parser.add_argument("--plain-steps", type=int, default=20,
                        help="events timed one call at a time in each plain form")
    parser.add_argument("--stream-events", type=int, default=256,
                        help="events of the serving stream (0: none)")
    parser.add_argument("--stream-rates", default="50,200,800,2000",
                        help="Poisson arrival rates (events/s) of the serving stream")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--cpu-fraction", type=float, default=0.02)
    parser.add_argument("--cpu-repeats", type=int, default=5)
    parser.add_argument("--no-counters", action="store_true",
                        help="skip the rocprofv3 counter passes (HBM traffic, engine busy)")
    parser.add_argument("--counter-dir", default=None,
                        help="keep the counter passes' CSV output in this folder")
    parser.add_argument("--profile-run", action="store_true",
                        help="only run two events of the plain route (counter passes)")
    parser.add_argument("--profile-stream", type=int, default=0,
                        help="only run this many serving-stream events of the plain route, "
                             "one call at a time (kernel traces of the stream)")
    parser.add_argument("--copy-kernel-mib", type=int, default=0,
                        help="only run a plain device copy of this size (HBM counter calibration)")
    args = parser.parse_args()
    if args.copy_kernel_mib:
        import torch
        a = torch.ones(args.copy_kernel_mib * 2 ** 18, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_kernel_mib, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(sharding the route by OM is a design note only)", file=sys.stderr)
        sys.exit(2)

    import warnings

    import torch
    from lighthit_amd import ballistic
    from lighthit_amd import order1 as order1_module
    from lighthit_amd.route import ResidentShower
    warnings.simplefilter("ignore", RuntimeWarning)

    if args.profile_stream:
        from lighthit_amd.synthetic import serving_stream
        workload, kernel, route = build_route(args.workload, args.precision, args.pack)
        stream = list(serving_stream(workload.detector, args.profile_stream))
        report = plain_stream(route, stream)["report"]
        print(json.dumps({"profile_stream": len(stream), **report}), flush=True)
        return

    if args.profile_run:
        workload, kernel, route = build_route(args.workload, args.precision, args.pack)
        for _ in range(2):
            route(ResidentShower(workload.elements))
        torch.cuda.synchronize()
        print(json.dumps({"profile_run": workload.name}), flush=True)
        return

    workload, kernel, route = build_route(args.workload, "fp32", args.pack)
    _, _, mixed = build_route(args.workload, "mixed", args.pack)
    served_route = mixed if args.precision == "mixed" else route
    elements = workload.elements
    print(f"[bench] {workload.name}: plain fp32, {args.plain_steps} events", file=sys.stderr,
          flush=True)
    plain_fp32 = time_plain(route, elements, args.plain_steps, args.warmup)
    print(f"[bench] plain mixed, {args.plain_steps} events", file=sys.stderr, flush=True)
    plain_mixed = time_plain(mixed, elements, args.plain_steps, args.warmup)
    plain_served_form = plain_mixed if args.precision == "mixed" else plain_fp32
    response = plain_served_form["response"]
    # work counters of the cost models over one untimed event (atomics)
    order1_module.count_work("cuda")
    ballistic.count_nodes("cuda")
    served_route(ResidentShower(elements))
    torch.cuda.synchronize()
    o1_counts = order1_module.work_counts("cuda")
    o0_nodes = ballistic.node_count()
    window = window_work(served_route, workload, response)
    order1_module.count_work("cuda", on=False)
    ballistic.count_nodes("cuda", on=False)
    precision = precision_report(workload, response, args.pack)

    stream = stream_plain = None
    rates = [float(r) for r in args.stream_rates.split(",") if r]
    if args.stream_events > 0:
        from lighthit_amd.synthetic import serving_stream
        print(f"[bench] serving stream: {args.stream_events} events, plain calls",
              file=sys.stderr, flush=True)
        stream = list(serving_stream(workload.detector, args.stream_events))
        stream_plain = plain_stream(served_route, stream)
    served = None
    if args.processes > 0:
        # the worker processes get the device memory this process cached
        from lighthit_amd import _lib as lib
        lib.release_workspaces()
        torch.cuda.empty_cache()
        print(f"[bench] serving {args.steps} x {args.processes} events from {args.processes} "
              f"processes ({args.precision})", file=sys.stderr, flush=True)
        served = serve_processes(args.workload, args.precision, workload.detector, elements,
                                 args.processes, args.steps, args.warmup, stream, stream_plain,
                                 rates, args.pack)
        results = served.pop("results")
        served["same_as_alone"] = bool(all(
            np.array_equal(r.charge_components_pe, response.charge_components_pe)
            and np.array_equal(r.components_pe, response.components_pe) for r in results))
    else:
        # plain form only (profiling runs): one step = one event
        served = {"events_per_s": plain_served_form["events_per_s"],
                  "ms_per_step": plain_served_form["ms_per_event"],
                  "latency_ms_poisson": plain_served_form["latency_ms"]}

    kernels = plain_served_form["kernel_gpu_ms"]
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_event"]) if kernels else None

    def entry(name):
        k = kernels[name]
        r = roofline_of(name, k["ms_per_launch"], workload, response, served_route,
                        counts=o1_counts, launches=k["launches_per_event"], nodes=o0_nodes,
                        window=window)
        if r is None:
            r = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": None, "algorithmic": "no cost model for this kernel"}
        return dict(r, kernel=name, ms_per_launch=k["ms_per_launch"])

    roofline = entry(dominant) if dominant else None
    stages = {stage: entry(name) for stage, name in STAGE_KERNELS.items() if name in kernels}
    if roofline is not None:
        counters, problem = (None, "not measured (--no-counters)")
        if not args.no_counters:
            counters, problem = counter_passes(args.workload, args.precision, dominant,
                                               args.counter_dir, args.pack)
        roofline["traffic"] = counters["traffic"] if counters else None
        roofline["traffic_note"] = counters["traffic_note"] if counters else problem
        if counters:
            for stage, figures in counters["engines"].items():
                if stage in stages:
                    stages[stage]["counters"] = figures
                elif stage == "dominant":
                    roofline["counters"] = figures
    print(json.dumps({"gpu_part": {"roofline": roofline, "stages": stages}}), file=sys.stderr,
          flush=True)
    baseline = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        baseline = cpu_baseline(workload, kernel, served_route, response, args.cpu_fraction,
                                args.cpu_repeats)
    for form in (plain_fp32, plain_mixed):
        form.pop("response")
    value = served["events_per_s"]
    line = {
        "metric": "events/sec + p50/p99 latency at <=1e-3 charge error vs FP64; served vs plain GPU",
        "value": value, "unit": "events/s", "n_gpus": 1,
        "steps": args.steps if args.processes > 0 else args.plain_steps,
        "warmup": args.warmup, "ms_per_step": served["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16/fp16 hi+lo MFMA, fp32 accumulate; fp32/fp64 vector" if
                 args.precision == "mixed" else "fp32/fp64",
        "data": "synthetic",
        "config": {"workload": workload.name, "order2_pack": args.pack,
                   "form": (f"served: {args.processes} worker processes on one GPU, each "
                            f"calling ShowerTransport(kernel, data)(shower) with precision "
                            f"'{args.precision}'; a step is a batch of {args.processes} events, "
                            "sources resident in HBM") if args.processes > 0 else
                           (f"plain: one ShowerTransport call at a time, precision "
                            f"'{args.precision}', sources resident in HBM"),
                   "events_per_step": max(1, args.processes),
                   "oms": len(workload.detector), "elements": len(elements),
                   "wavelength_nodes": len(workload.wavelength.wavelength_nm),
                   "bins": len(workload.edges_ns) - 1,
                   "active_oms": int(np.count_nonzero(response.active)),
                   "order1_oms": int(np.count_nonzero(
                       np.asarray(response.metadata["prompt_order1_level"]) > 0)),
                   "charge_pe_by_order": [float(x) for x in
                                          response.charge_components_pe.sum(axis=0)],
                   "latency_ms": dict(served["latency_ms_poisson"],
                                      form="served, Poisson arrivals at 70% of throughput"
                                      if args.processes > 0 else "plain, one call at a time"),
                   "error_vs_fp64": precision,
                   "served": served,
                   "plain_fp32": plain_fp32, "plain_mixed": plain_mixed,
                   "served_over_plain_fp32": value / plain_fp32["events_per_s"],
                   "served_over_plain_mixed": value / plain_mixed["events_per_s"],
                   "host_stage_ms": {k: 1e3 * v for k, v in
                                     response.metadata["timings"].items()},
                   "stage_rooflines": stages,
                   "work_counts": {"order0_nodes": o0_nodes, "order1": o1_counts,
                                   "order2_window": window}},
        "roofline": roofline, "cpu_baseline": baseline,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
