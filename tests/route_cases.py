"""The demonstration route: kernel, synthetic order >=2 pack and route object."""
import numpy as np


def demonstration_route(device="cuda", tables=False):
    """The demonstration event's route; ``tables`` builds the order >=2 pack from
    the RTE tables (lighthit_amd.tables) instead of the seeded synthetic pack."""
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import RouteConfig, RouteKernel, ShowerTransport
    from lighthit_amd.synthetic import plain_small, synthetic_pack
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    pack = KernelPack()
    if tables:
        from lighthit_amd.tables import build_pack
        pack.add(*build_pack(kernel))
    else:
        pack.add(*synthetic_pack(w.medium, kernel.wavelength))
    route = ShowerTransport(kernel, pack, device=device)
    return w, kernel, pack, route


def near_event(om=20, elements=400, impact=(1.0, 0.3, 0.0), seed=5):
    """A short shower passing about 1 m from OM ``om`` of the demonstration array
    (inside the tables' 2 m radial minimum, outside the 0.2 m glass)."""
    from dataclasses import replace

    from lighthit_amd.sources import SourcePose, SyntheticShower
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    position = w.detector.positions_m[om]
    direction = np.array([0.05, 0.0, 1.0])
    pose = SourcePose(position + np.asarray(impact), direction / np.linalg.norm(direction))
    shower = SyntheticShower.gaussian(pose, charged_track_length_m=30.0, elements=elements,
                                      seed=seed)
    return w, replace(shower.elements, cone_model="spectral")


def demonstration_route_only(device="cuda"):
    """The demonstration route alone (a picklable factory for worker processes)."""
    return demonstration_route(device)[3]


def detector_route_only(precision="fp32", device="cuda:0"):
    """The detector-scale route in ``precision`` (a picklable factory for worker
    processes)."""
    from dataclasses import replace

    from lighthit_amd.route import ShowerTransport
    from lighthit_amd.synthetic import workload_route
    workload, kernel, route = workload_route("detector", device=device)
    if precision != route.config.precision:
        route = ShowerTransport(route.kernel, route.data,
                                config=replace(route.config, precision=precision))
    return route


def _joined(first, second):
    from dataclasses import replace
    names = ("start_m", "direction", "length_m", "coefficient0", "coefficient2", "beta",
             "start_ns", "end_ns", "row_index", "uid")
    return replace(first, **{name: np.concatenate([getattr(first, name), getattr(second, name)])
                             for name in names})


def small_shower(elements=300, seed=7, offset=(-12.0, 8.0, 0.0),
                 direction=(0.35, -0.15, 0.925), track_m=300.0):
    """A demonstration-like shower with fewer elements (spectral cones)."""
    from dataclasses import replace

    from lighthit_amd.sources import SourcePose, SyntheticShower
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    centre = w.detector.positions_m.mean(axis=0)
    pose = SourcePose(centre + np.asarray(offset, float), np.asarray(direction, float))
    shower = SyntheticShower.gaussian(pose, charged_track_length_m=track_m, elements=elements,
                                      seed=seed)
    return replace(shower.elements, cone_model="spectral")


def delayed_pair(elements=300, delay_ns=250.0):
    """A shower and the same shower with a copy emitted ``delay_ns`` later (the
    reference's delayed-light invariant, tests/test_shower_route.py:200-228)."""
    from dataclasses import replace
    first = small_shower(elements)
    late = replace(first, start_ns=first.start_ns + delay_ns, end_ns=first.end_ns + delay_ns,
                   row_index=first.row_index + len(first))
    return first, _joined(first, late)


def long_source(elements=300, separation_m=200.0):
    """Two showers ``separation_m`` apart along the array's x axis: the source extent
    exceeds the 150 m radial range, so the route takes per-OM (local) time origins."""
    a = small_shower(elements, seed=7, offset=(-separation_m / 2, 8.0, 0.0))
    b = small_shower(elements, seed=8, offset=(separation_m / 2, -8.0, 10.0),
                     direction=(-0.30, 0.20, 0.93))
    return _joined(a, b)
