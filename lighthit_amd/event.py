"""Events made of several shower sources: one clock, one OM selection.

Mirror of the event composer of the reference (``event.py:1-301``) for shower
components, the caller of the shower route (``EventTransport.shower_route``,
``event.py:124-151``). Light adds, so an event's response is the sum of its
sources' responses, with two steps that cannot be done source by source:

* one time origin per OM for every source (a lower bound on the earliest
  arrival from any element of the event at the fastest group speed), so the
  bins of different sources add on one clock;
* one OM selection: an OM is active when the summed charge reaches the
  threshold. Each source is first run on the whole detector with the charge
  of the sources already run added to its selection (``selection_extra_
  charge``); then, while some OM could still reach the threshold
  (``q0 + safety * (|q1| + |q2|)`` of the sources that only estimated it, the
  detailed charge of the others) and a source lacks its detail there, that
  source is run again on exactly those OMs (``receivers=``).

The element bounds behind the time origins run on the GPU
(``lh_segment_bounds``); every source runs through
:class:`lighthit_amd.route.ShowerTransport`. Muon components need the
reference's muon route, which this package does not provide: they are
refused with a ``ValueError``, as are calibration flashes.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter, process_time

import numpy as np

__all__ = ["SourceComponent", "EventTransport", "common_time_origins"]

C_M_PER_NS = 0.299792458


@dataclass(frozen=True)
class SourceComponent:
    """One light source of an event (``event.py:41-62``): ``kind`` "shower" (or
    "muon", with its ``ancestry`` and ``primary_uid``), its spectral elements
    in ``source`` and a free-text ``label``."""
    kind: str
    source: object
    ancestry: dict | None = None
    primary_uid: int | None = None
    label: str = ""

    def __post_init__(self):
        if self.kind not in ("shower", "muon"):
            raise ValueError("component kind must be 'shower' or 'muon'")
        if self.kind == "muon" and (self.ancestry is None or self.primary_uid is None):
            raise ValueError("a muon component needs ancestry and primary_uid")
        elements = getattr(self.source, "elements", self.source)
        if getattr(elements, "cone_model", None) != "spectral":
            raise ValueError("components require spectral cones (cone_model='spectral')")


def _elements(source):
    return getattr(source, "elements", source)


def common_time_origins(kernel, sources, device="cuda"):
    """Per-OM lower bound on the earliest arrival from any element of
    ``sources`` (``event.py:65-81``): start time plus the distance to the
    element at the fastest group speed of the kernel's wavelength nodes, the
    minimum over every element of every source (``lh_segment_bounds``)."""
    from .route import segment_bounds
    group = float(np.min(kernel.medium.sample(
        np.asarray(kernel.wavelength.wavelength_nm, float))["group_index"]))
    positions = np.ascontiguousarray(kernel.detector.positions_m, float)
    origins = np.full(len(positions), np.inf)
    for source in sources:
        _, earliest = segment_bounds(positions, _elements(source), group / C_M_PER_NS, device)
        origins = np.minimum(origins, earliest)
    return origins


@dataclass
class _Part:
    """What one component has contributed so far."""
    component: SourceComponent
    charge: np.ndarray          # (N, 3)
    bins: np.ndarray            # (N, bins, 3), rows of detailed OMs
    spectrum: np.ndarray        # (frequencies, N)
    detailed: np.ndarray        # (N,) bool: charge and bins computed in detail
    runs: list = field(default_factory=list)


class EventTransport:
    """All orders for an event given as a list of shower components
    (``event.py:94-301``). ``kernel`` and ``shower_data`` build the shower route
    (or pass ``route``, a ready :class:`ShowerTransport`); ``safety`` inflates
    the scattered charges a route only estimated when deciding which OMs must
    be completed in every source; ``max_rounds`` bounds the completion rounds."""

    def __init__(self, kernel, *, shower_data=None, route=None, safety=3.0, max_rounds=4,
                 device="cuda"):
        if not np.isfinite(safety) or safety < 1 or int(max_rounds) < 1:
            raise ValueError("safety must be >= 1 and max_rounds positive")
        self.kernel = kernel
        self.shower_data = shower_data
        self.safety, self.max_rounds = float(safety), int(max_rounds)
        self.device = device
        self._shower = route

    def shower_route(self):
        """The :class:`lighthit_amd.route.ShowerTransport` of this detector."""
        if self._shower is None:
            from .route import ShowerTransport
            self._shower = ShowerTransport(self.kernel, self.shower_data, device=self.device)
        return self._shower

    def muon_route(self):
        raise ValueError("muon components need the reference's muon route, which this package "
                         "does not provide")

    def _run(self, component, **options):
        """One route call of one component (whole detector, or ``receivers``)."""
        if component.kind != "shower":
            self.muon_route()
        return self.shower_route()(component.source, **options)

    def __call__(self, components, *, progress=None):
        """Response (``HybridTransportResponse``) of the event's components."""
        if not isinstance(components, (list, tuple)):
            raise ValueError("an event is a list of SourceComponent (calibration flashes need "
                             "the reference's laser kernel)")
        components = list(components)
        if not components:
            raise ValueError("an event needs at least one source component")
        if not all(isinstance(c, SourceComponent) for c in components):
            raise ValueError("an event is a list of SourceComponent")
        if len(components) == 1:
            response = self._run(components[0])
            response.metadata["event_components"] = [self._describe(components[0])]
            return response
        return self._compose(components, progress)

    @staticmethod
    def _describe(component):
        return dict(kind=component.kind, label=component.label,
                    elements=len(_elements(component.source)),
                    primary_uid=component.primary_uid)

    def _compose(self, components, progress):
        from .route import HybridTransportResponse
        began, cpu = perf_counter(), process_time()
        kernel = self.shower_route().kernel
        n = len(kernel.detector.positions_m)
        edges = np.asarray(kernel.config.relative_time_edges_ns, float)
        threshold_pe = float(kernel.config.threshold_pe)
        detail = getattr(kernel.config, "detail_threshold_pe", None)
        # the completion rounds decide at the lower of the two thresholds
        threshold = threshold_pe if detail is None else min(threshold_pe, float(detail))
        for component in components:
            if component.kind != "shower":
                self.muon_route()
        origins = common_time_origins(kernel, [c.source for c in components], self.device)
        parts, known, omega = [], np.zeros(n), None
        for component in components:
            started = process_time()
            response = self._run(component, time_origin_ns=origins, selection_extra_charge=known)
            if omega is None:
                omega = np.asarray(response.omega_per_ns)
            elif not np.allclose(response.omega_per_ns, omega, rtol=0, atol=1e-12):
                raise ValueError("source routes use different RTE frequency grids")
            if not np.allclose(response.time_origin_ns, origins, rtol=0, atol=1e-9):
                raise ValueError("a source route did not use the event clock")
            charge = np.asarray(response.charge_components_pe, float).copy()
            # computed in detail: the route's active OMs, and OMs the source
            # does not reach at all (exactly zero charge: nothing to complete)
            detailed = np.asarray(response.active, bool) | np.all(charge == 0.0, axis=1)
            part = _Part(component, charge,
                         np.where(detailed[:, None, None], response.components_pe, 0.0),
                         np.asarray(response.spectrum_ge2_pe).copy(), detailed)
            part.runs.append(dict(receivers="all", cpu_seconds=process_time() - started,
                                  route_active=int(np.count_nonzero(response.active))))
            parts.append(part)
            known = known + charge.sum(axis=1)
            if progress:
                progress("component", self._describe(component))
        rounds = []
        for number in range(self.max_rounds):
            bound = np.zeros(n)
            for part in parts:
                q = part.charge
                estimate = np.maximum(q[:, 0], 0.0) + self.safety * np.abs(q[:, 1:]).sum(axis=1)
                bound += np.where(part.detailed, q.sum(axis=1), estimate)
            wanted = bound >= threshold
            record = dict(round=number, candidates=int(wanted.sum()), recomputed={})
            for part in parts:
                need = np.flatnonzero(wanted & ~part.detailed
                                      & (np.abs(part.charge).sum(axis=1) > 0))
                if not len(need):
                    continue
                started = process_time()
                local = self._run(part.component, time_origin_ns=origins[need], receivers=need)
                part.charge[need] = local.charge_components_pe
                part.bins[need] = local.components_pe
                part.spectrum[:, need] = local.spectrum_ge2_pe
                part.detailed[need] = True
                part.runs.append(dict(receivers=need.tolist(),
                                      cpu_seconds=process_time() - started))
                record["recomputed"][part.component.label or part.component.kind] = len(need)
            rounds.append(record)
            if not record["recomputed"]:
                break
        else:
            raise RuntimeError("source composition did not converge; raise max_rounds")
        charge = sum(part.charge for part in parts)
        detailed = np.logical_and.reduce([part.detailed for part in parts])
        total = charge.sum(axis=1)
        if np.any((total >= threshold) & ~detailed):
            raise RuntimeError("an OM above threshold lacks the detailed charge of a source")
        active = (total >= threshold) & detailed
        if threshold == 0:
            active = detailed & (np.abs(charge).sum(axis=1) > 0)
        components_pe = np.zeros((n, len(edges) - 1, 3))
        components_pe[active] = sum(part.bins for part in parts)[active]
        spectrum = sum(part.spectrum for part in parts)
        if detail is not None and threshold < threshold_pe:
            # every decision at the detail threshold; ``active`` at the event's own
            active = active & (total >= threshold_pe)
        metadata = dict(
            route="lighthit.event", elapsed_seconds=perf_counter() - began,
            cpu_seconds=process_time() - cpu, threshold_pe=threshold_pe,
            time_origin="earliest arrival from any element of the event, fastest group speed",
            selection="summed charge of all sources; OMs completed in every source while "
                      "q0 + safety*(|q1|+|q2|) of estimated sources can reach the threshold",
            safety=self.safety, completion_rounds=rounds, active_modules=int(active.sum()),
            event_components=[dict(self._describe(part.component), runs=part.runs,
                                   charge_pe=float(part.charge.sum()),
                                   detailed_modules=int(part.detailed.sum()))
                              for part in parts],
            inactive_charge_may_be_estimated=bool(np.any(~detailed)),
            computed_orders=[0, 1, ">=2"])
        return HybridTransportResponse(kernel.detector, omega, edges, spectrum, components_pe,
                                       charge, origins, active, "event", metadata)
