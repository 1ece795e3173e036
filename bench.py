"""Benchmark of the ShowerTransport call on one MI355X.

One step = one shower event through ``ShowerTransport(kernel, data)(shower)``:
the whole route of the plain FP32 GPU form, as a user calls it — node-band
order >=2 compile, radial windows, collapsed predictor, orders 0 and 1, exact
zero-frequency order >=2 where the threshold needs it, the full 65-frequency
apply at the active OMs and the readout into time bins. The detector
description, the RTE tables and the folded order >=2 pack are resident before
the timed region (built once, as a served route keeps them); every step hands
the route a fresh source object, so the per-event compile is never a cache hit.

Workload (default): the detector-scale event — 2304 OMs (8 clusters x 8
strings x 36 OMs), a 50,000-element Gaussian shower of seed 11 with a 1500 m
charged track, 8 wavelength nodes, 1 ns bins over -60..745 ns, 65 RTE and 321
readout frequencies, degree 8, azimuthal degree 2, radial window 2..150 m,
order-1 selection radius 300 m. The order >=2 pack is the seeded synthetic
pack of lighthit_amd.synthetic (the reference's pack needs its RTE solver).

Prints one JSON line: events/s, p50/p99 latency, GPU time per stage and per
dominant kernel (HIP events on the launch stream), the roofline of the
dominant kernel and the FP64 oracle timed on the host over a bounded sample.
``--gpus`` above 1 exits: this build runs on one GPU.
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
O1_MEMBER_BYTES = 22 * 8 + 4

# FP64 operations per unit of the order-1 batch kernel (each add, multiply,
# divide, square root or exponential counted once):
#   profile term (one path node x one group): scattering geometry 16, the two
#     cubic acceptance polynomials 28, Henyey-Greenstein weight 8, sums 4;
#   time panel: quadratic coefficients 30, two attenuation factors 4, the
#     exponential-quadratic integral 154 (three 24-term Taylor series and the
#     combination);
#   bin event: one more exponential-quadratic integral over a bin piece 160.
O1_PROFILE_TERM_FLOPS = 56
O1_PANEL_FLOPS = 188
O1_EVENT_FLOPS = 160

#: Kernels the library times with HIP events on their stream.
TIMED_KERNELS = ("order0_pairs<false>", "order0_pairs<true>", "o1_batch_profile",
                 "o1_batch_deposit", "o1_batches", "o1_group_sums", "ge2_wide", "ge2_narrow",
                 "readout_kernel", "compile_deposit", "compile_phase")


def build_route(name):
    """Workload, route (tables resident) and the event's source."""
    from lighthit_amd.synthetic import workload_route
    return workload_route(name)


def route_factory(name):
    """The route alone, built the same way in a serving worker process."""
    return build_route(name)[2]


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


def cpu_baseline(workload, kernel, route, response, fraction):
    """The FP64 oracle on the host cores, one thread, over a bounded sample of
    the same event: order 0 (charge at the selected OMs, bins at the active
    OMs) on every ``1/fraction``-th element scaled to all elements; order 1,
    the node-band compile, the exact zero-frequency and full order >=2 apply
    and the readout in full on the OM sets the GPU route chose."""
    from threadpoolctl import threadpool_limits

    from lighthit_amd.axial import AxisFrame
    from lighthit_amd.optics import prepare_optics
    from oracle import ge2 as oracle_ge2
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
    optics = prepare_optics(kernel.medium, d)
    radii = d.receiver_surface.radii(len(d))
    level1 = np.asarray(response.metadata["prompt_order1_level"])
    level2 = np.asarray(response.metadata["ge2_charge_level"])
    work1 = np.flatnonzero(level1 > 0)
    candidates = np.flatnonzero(level2 == 2)
    active = np.flatnonzero(response.active)
    full = np.flatnonzero(response.active & (level2 > 0))
    origins = np.asarray(response.time_origin_ns, float)
    edges = np.asarray(kernel.config.relative_time_edges_ns, float)
    omega = np.asarray(kernel.config.omega_per_ns, float)
    stride = max(1, int(round(1 / fraction)))
    keep = np.arange(len(elements)) % stride == 0
    sample = elements.subset(keep)
    seconds = {}
    with threadpool_limits(limits=1):
        began = time.perf_counter()
        oracle_order0.first_entry(d.positions_m[work1], (-d.orientations)[work1],
                                  d.effective_area_m2[work1], radii[work1], sample, optics,
                                  alpha, rel_tol=2e-6)
        if len(active):
            oracle_order0.first_entry(d.positions_m[active], (-d.orientations)[active],
                                      d.effective_area_m2[active], radii[active], sample,
                                      optics, alpha, origins=origins[active], edges=edges,
                                      rel_tol=2e-6)
        seconds["order0"] = (time.perf_counter() - began) * len(elements) / len(sample)
        began = time.perf_counter()
        spectral = oracle_order1.spectral_tables(kernel.medium, d, kernel.wavelength)
        charge = response.charge_components_pe
        oracle_order1.spectral_order1(elements, d.positions_m, (-d.orientations), d.effective_area_m2,
                                      origins, work1, spectral, alpha, float(kernel.medium.g),
                                      edges, float(kernel.config.threshold_pe), charge[:, 0],
                                      2.0 * np.abs(charge[:, 2]), oracle_order1.PromptSettings())
        seconds["order1"] = time.perf_counter() - began
        began = time.perf_counter()
        wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
        max_phase = float(np.max(kernel.medium.sample(wavelengths)["phase_index"]))
        frame = AxisFrame.of(elements.field(0, phase_index=max_phase))
        sources = oracle_route._node_sources(
            elements, wavelengths, fold, omega, frame, int(kernel.config.source_degree),
            int(kernel.config.azimuthal_degree), route.config.basis)
        seconds["compile"] = time.perf_counter() - began
        began = time.perf_counter()
        window = tuple(kernel.config.radial_range_m)
        pos, looks, area = d.positions_m, (-d.orientations), d.effective_area_m2
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
    total = sum(seconds.values())
    return {"value": 1.0 / total, "unit": "events/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 oracle (C and numpy, one thread), one event: order 0 on every "
                       f"{stride}th element ({len(sample)} of {len(elements)}) at the reference "
                       f"tolerance 2e-6, scaled to all elements, at {len(work1)} selected and "
                       f"{len(active)} active OMs; order 1 on {len(work1)} OMs, node-band "
                       f"compile, order >=2 zero frequency at {len(candidates)} and full "
                       f"apply and readout at {len(full)} OMs in full; windows and predictor "
                       f"not timed"),
            "stage_seconds": seconds}


# --------------------------------------------------------- HBM traffic (pmc)

def _counter_kib(db_dir, counter, name_part):
    """Average KiB per dispatch of ``counter`` for kernels whose name holds ``name_part``."""
    import glob
    import sqlite3
    values = []
    for db in glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True):
        cur = sqlite3.connect(db).cursor()
        try:
            rows = cur.execute("select kernel_name, value from counters_collection "
                               "where counter_name=?", (counter,)).fetchall()
        except sqlite3.Error:
            continue
        values += [float(v) for k, v in rows if name_part in k]
    return sum(values) / len(values) if values else None


def measure_traffic(workload_flag, kernel_name, keep_dir=None):
    """HBM bytes per launch of ``kernel_name`` from two counter runs, each
    scaled by what the same counter reads for a plain copy of known size.
    ``keep_dir`` keeps the counter databases there instead of a temporary folder."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    if keep_dir:
        base = os.path.abspath(keep_dir)
        os.makedirs(base, exist_ok=True)
    else:
        base = tempfile.mkdtemp(prefix="lh_traffic_")
    me = os.path.abspath(__file__)
    runs = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for kind, args in (("bench", ["--steps", "1", "--warmup", "1", "--no-cpu-baseline",
                                      "--no-traffic", "--workers", "1", "--processes", "0",
                                      "--workload",
                                      workload_flag]),
                           ("copy", ["--copy-kernel-mib", "256"])):
            out = os.path.join(base, f"{kind}_{counter}")
            cmd = ["rocprofv3", "--pmc", counter, "-d", out, "-o", "run", "--",
                   sys.executable, me, *args]
            print(f"[bench] counter pass {counter} {kind}", file=sys.stderr, flush=True)
            done = subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd="/tmp",
                                  env=dict(os.environ, TMPDIR="/tmp"))
            if done.returncode != 0:
                return None, f"rocprofv3 {counter} {kind} failed ({done.returncode})"
            runs[(kind, counter)] = out
    result = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        copy_kib = _counter_kib(runs[("copy", counter)], counter, "copyBuffer")
        kernel_kib = _counter_kib(runs[("bench", counter)], counter, kernel_name)
        if not copy_kib or kernel_kib is None:
            return None, f"no {counter} samples"
        scale = 256 * 1024 / copy_kib    # the copy moves 256 MiB each way
        result[counter] = (kernel_kib * scale * 1024, scale, kernel_kib)
    traffic = result["FETCH_SIZE"][0] + result["WRITE_SIZE"][0]
    note = (f"{kernel_name}: FETCH_SIZE {result['FETCH_SIZE'][2]:.0f} KiB x"
            f"{result['FETCH_SIZE'][1]:.3f} + WRITE_SIZE {result['WRITE_SIZE'][2]:.0f} KiB x"
            f"{result['WRITE_SIZE'][1]:.3f} per dispatch (copy-calibrated), bytes per launch")
    return traffic, note


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

def roofline_of(name, kernel_ms, workload, response, route, counts=None, launches=1):
    """Roofline entry of kernel ``name`` at its average launch duration."""
    d = workload.detector
    n_el = len(workload.elements)
    if name.startswith("order0_pairs"):
        bins = 0 if name.endswith("<false>)") or name.endswith("<false>") else (
            len(workload.edges_ns) - 1)
        oms = len(d) if bins == 0 else int(np.count_nonzero(response.active))
        achieved = order0_bytes(n_el, oms, bins) / (kernel_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "algorithmic": f"48 B/element x {n_el} + 36 B/OM x {oms} + 8 B x {max(1, bins)} "
                               f"outputs/OM"}
    if name in ("o1_batch_deposit", "o1_batch_profile") and counts is not None:
        terms, panels, events = (c / launches for c in counts[:3])
        if name == "o1_batch_profile":
            flops = O1_PROFILE_TERM_FLOPS * terms
            algorithmic = (f"per launch {terms:.4g} profile terms x {O1_PROFILE_TERM_FLOPS} "
                           f"FP64 flops (counted in the kernel)")
        else:
            flops = O1_PANEL_FLOPS * panels + O1_EVENT_FLOPS * events
            algorithmic = (f"per launch {panels:.4g} panels x {O1_PANEL_FLOPS} + {events:.4g} "
                           f"bin events x {O1_EVENT_FLOPS} FP64 flops (counted in the kernel)")
        achieved = flops / (kernel_ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "algorithmic": algorithmic,
                "note": "FP64 vector work (no matrix form); priced against the FP64 peak"}
    if name == "o1_group_sums" and counts is not None and len(counts) > 3:
        members = counts[3] / launches
        achieved = members * O1_MEMBER_BYTES / (kernel_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "algorithmic": f"per launch {members:.4g} group members x {O1_MEMBER_BYTES} B "
                               "(the member's pencil fields, 22 FP64 values, and its 4-byte "
                               "index; counted in the kernel)",
                "note": "gathers in key order; the FP64 geometry per member is not priced"}
    if name == "o1_batches" and counts is not None:
        terms, panels, events = (c / launches for c in counts[:3])
        flops = (O1_PROFILE_TERM_FLOPS * terms + O1_PANEL_FLOPS * panels
                 + O1_EVENT_FLOPS * events)
        achieved = flops / (kernel_ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "algorithmic": f"per launch {terms:.4g} profile terms x {O1_PROFILE_TERM_FLOPS} "
                               f"+ {panels:.4g} panels x {O1_PANEL_FLOPS} + {events:.4g} bin "
                               f"events x {O1_EVENT_FLOPS} FP64 flops (counted in the kernel)",
                "note": "FP64 vector work (no matrix form); priced against the FP64 peak"}
    if name == "ge2_wide":
        from lighthit_amd.ge2 import canonical_pairs
        level2 = np.asarray(response.metadata["ge2_charge_level"])
        full = np.flatnonzero(response.active & (level2 > 0))
        terms = ge2_window_terms(route, d.positions_m[full], workload.radial_range_m)
        pack = route.tables()["device_pack"]
        per = ge2_flops_per_term(workload.source_degree, workload.azimuthal_degree,
                                 pack.acceptance_degree,
                                 len(canonical_pairs(pack.acceptance_degree)[0]))
        flops = float(terms) * len(workload.omega_per_ns) * per
        achieved = flops / (kernel_ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "algorithmic": f"{terms} (OM, point, band) terms in the window x "
                               f"{len(workload.omega_per_ns)} frequencies x {per} flops"}
    return None


def main():
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", choices=["detector", "small", "heavy"], default="detector")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--cpu-fraction", type=float, default=0.05)
    parser.add_argument("--workers", type=int, default=3,
                        help="concurrent route calls of the served form (1: plain only)")
    parser.add_argument("--served-events", type=int, default=120)
    parser.add_argument("--stream-events", type=int, default=256,
                        help="events of the serving stream (0: none)")
    parser.add_argument("--stream-rates", default="50,200,800,2000",
                        help="Poisson arrival rates (events/s) of the serving stream")
    parser.add_argument("--processes", type=int, default=8,
                        help="also serve from this many worker processes (0: threads only)")
    parser.add_argument("--no-traffic", action="store_true",
                        help="skip the rocprofv3 counter runs that measure HBM traffic")
    parser.add_argument("--traffic-dir", default=None,
                        help="keep the counter runs' databases in this folder")
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
    from lighthit_amd._lib import KernelTimer, StageClock
    from lighthit_amd.route import ResidentShower
    from lighthit_amd.serve import ShowerServer, latency_summary

    workload, kernel, route = build_route(args.workload)
    elements = workload.elements
    route.tables()
    print(f"[bench] {workload.name}: warmup {args.warmup}", file=sys.stderr, flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        first = None
        for _ in range(args.warmup):
            first = route(ResidentShower(elements))
        # every step's source is uploaded before the timed region (a fresh
        # object per step, so the per-event compile always runs)
        resident = [ResidentShower(elements) for _ in range(args.steps)]
        torch.cuda.synchronize()
        latencies = []
        same = True
        print(f"[bench] timing {args.steps} events (plain, one at a time)", file=sys.stderr,
              flush=True)
        from lighthit_amd import order1 as order1_module
        # the order-1 work counters (for the cost models) over one untimed
        # event: counting adds atomics, so it stays out of the timed calls
        order1_module.count_work("cuda")
        route(ResidentShower(elements))
        torch.cuda.synchronize()
        o1_counts = order1_module.work_counts("cuda")
        counted_events = 1
        order1_module.count_work("cuda", on=False)
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
                                                first.charge_components_pe))
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - began
            stage_ms = {k: v / args.steps for k, v in clock.milliseconds().items()}
            kernels = {}
            for name in TIMED_KERNELS:
                total, launches = timer.read(name)
                if launches:
                    kernels[name] = {"ms_per_launch": total / launches,
                                     "launches_per_event": launches / args.steps,
                                     "ms_per_event": total / args.steps}
            timer.reset()
        del resident
        # the same calls with the host arrays handed over (PCIe upload inside)
        host_began = time.perf_counter()
        for _ in range(args.steps):
            route(copy.copy(elements))
        torch.cuda.synchronize()
        host_elapsed = time.perf_counter() - host_began
        served = None
        if args.workers > 1:
            print(f"[bench] serving {args.served_events} events on {args.workers} workers",
                  file=sys.stderr, flush=True)
            with ShowerServer(route, workers=args.workers) as server:
                server.map([ResidentShower(elements) for _ in range(2 * args.workers)])
                batch = [ResidentShower(elements) for _ in range(args.served_events)]
                torch.cuda.synchronize()
                served_began = time.perf_counter()
                futures = [server.submit(source) for source in batch]
                results = [f.result() for f in futures]
                torch.cuda.synchronize()
                served_elapsed = time.perf_counter() - served_began
            mixed_same = all(np.array_equal(r.charge_components_pe, first.charge_components_pe)
                             and np.array_equal(r.components_pe, first.components_pe)
                             for r in results)
            served = {"events_per_s": len(batch) / served_elapsed, "events": len(batch),
                      "workers": args.workers,
                      "ms_per_event": 1e3 * served_elapsed / len(batch),
                      "latency_ms_all_submitted_at_once": latency_summary(futures),
                      "same_as_alone": bool(mixed_same)}
    stream_plain = None
    if args.processes > 0 and args.stream_events > 0:
        from lighthit_amd.synthetic import serving_stream
        print(f"[bench] serving stream: {args.stream_events} events, plain calls",
              file=sys.stderr, flush=True)
        stream = list(serving_stream(workload.detector, args.stream_events))
        stream_plain = plain_stream(route, stream)
        # the worker processes get the device memory this process cached
        from lighthit_amd import _lib as lib
        lib.release_workspaces()
        torch.cuda.empty_cache()
    processes = None
    if args.processes > 0:
        import functools

        from lighthit_amd.serve import ProcessShowerServer
        print(f"[bench] serving {args.served_events} events from {args.processes} processes",
              file=sys.stderr, flush=True)
        with ProcessShowerServer(functools.partial(route_factory, args.workload),
                                 processes=args.processes, detector=workload.detector,
                                 device="cuda:0") as server:
            server.map([elements for _ in range(2 * args.processes)])
            # every event's source is staged in device memory in the workers
            # before the timed region (a separate source per event)
            staged = [server.stage(elements) for _ in range(args.served_events)]
            began_p = time.perf_counter()
            futures = [server.submit(handle) for handle in staged]
            results = [f.result() for f in futures]
            elapsed_p = time.perf_counter() - began_p
            for handle in staged:
                server.release(handle)
            # latency under Poisson arrivals at 70% of the throughput just measured
            rate = 0.7 * len(results) / elapsed_p
            staged = [server.stage(elements) for _ in range(args.served_events)]
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
            if stream_plain is not None:
                rates = [float(r) for r in args.stream_rates.split(",") if r]
                stream_report = served_stream(server, stream, stream_plain, rates)
        same_p = all(np.array_equal(r.charge_components_pe, first.charge_components_pe)
                     and np.array_equal(r.components_pe, first.components_pe)
                     for r in results)
        processes = {"events_per_s": len(results) / elapsed_p, "events": len(results),
                     "processes": args.processes,
                     "ms_per_event": 1e3 * elapsed_p / len(results),
                     "latency_ms_all_submitted_at_once": latency_summary(futures),
                     "latency_ms_poisson": dict(latency_summary(arrivals),
                                                rate_events_per_s=rate),
                     "same_as_alone": bool(same_p),
                     "sources": "staged in device memory in every worker before timing",
                     "stream": stream_report}
    ms_per_step = 1e3 * elapsed / args.steps
    lat = np.asarray(latencies) * 1e3
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_event"]) if kernels else None
    roofline = None
    if dominant is not None:
        roofline = roofline_of(dominant, kernels[dominant]["ms_per_launch"], workload, response,
                               route, counts=o1_counts,
                               launches=kernels[dominant]["launches_per_event"] * counted_events)
        if roofline is None:
            roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": None, "algorithmic": "no cost model for this kernel"}
        roofline["kernel"] = dominant
        roofline["ms_per_launch"] = kernels[dominant]["ms_per_launch"]
        traffic, note = (None, "not measured (--no-traffic)")
        if not args.no_traffic:
            traffic, note = measure_traffic(args.workload, dominant.split("<")[0],
                                             args.traffic_dir)
        roofline["traffic"] = traffic
        roofline["traffic_note"] = note
    print(json.dumps({"gpu_part": {"ms_per_step": ms_per_step, "kernels": kernels,
                                   "stages": stage_ms, "roofline": roofline}}),
          file=sys.stderr, flush=True)
    baseline = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        baseline = cpu_baseline(workload, kernel, route, response, args.cpu_fraction)
    plain = {"events_per_s": 1e3 / ms_per_step, "ms_per_event": ms_per_step,
             "latency_ms": {"p50": float(np.percentile(lat, 50)),
                            "p99": float(np.percentile(lat, 99)), "max": float(lat.max())},
             "events_per_s_host_arrays": args.steps / host_elapsed}
    if processes is not None and (served is None
                                  or processes["events_per_s"] >= served["events_per_s"]):
        value, step_ms, form = (processes["events_per_s"], processes["ms_per_event"],
                                f"served: {args.processes} worker processes, one "
                                "ShowerTransport call each at a time")
        value_latency = dict(processes["latency_ms_poisson"], form="served, Poisson arrivals")
    elif served is not None:
        value, step_ms, form = (served["events_per_s"], served["ms_per_event"],
                                f"served: {args.workers} concurrent ShowerTransport calls")
        value_latency = dict(served["latency_ms_all_submitted_at_once"],
                             form="served, all submitted at once")
    else:
        value, step_ms, form = plain["events_per_s"], ms_per_step, "plain: one call at a time"
        value_latency = dict(plain["latency_ms"], form="plain")
    line = {
        "metric": "events/sec + p50/p99 latency at <=1e-3 charge error vs FP64; served vs plain GPU",
        "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": workload.name,
                   "form": form + " of ShowerTransport(kernel, data)(shower), FP32/FP64 GPU "
                           "path, sources resident in HBM",
                   "oms": len(workload.detector), "elements": len(elements),
                   "wavelength_nodes": len(workload.wavelength.wavelength_nm),
                   "bins": len(workload.edges_ns) - 1,
                   "active_oms": int(np.count_nonzero(response.active)),
                   "order1_oms": int(np.count_nonzero(
                       np.asarray(response.metadata["prompt_order1_level"]) > 0)),
                   "charge_pe_by_order": [float(x) for x in
                                          response.charge_components_pe.sum(axis=0)],
                   "plain": plain, "served": served, "served_processes": processes,
                   "served_over_plain": value / plain["events_per_s"],
                   "latency_ms": value_latency,
                   "repeatable": same,
                   "stage_gpu_ms": stage_ms, "kernel_gpu_ms": kernels,
                   "host_stage_ms": {k: 1e3 * v for k, v in
                                     response.metadata["timings"].items()},
                   "host_prompt_ms": {k: 1e3 * v for k, v in
                                      response.metadata["prompt_timings"].items()}},
        "roofline": roofline, "cpu_baseline": baseline,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
