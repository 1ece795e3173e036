"""The demonstration route: kernel, synthetic order >=2 pack and route object."""
import numpy as np


def demonstration_route(device="cuda"):
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import RouteConfig, RouteKernel, ShowerTransport
    from lighthit_amd.synthetic import plain_small, synthetic_pack
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    pack = KernelPack()
    key = pack.add(*synthetic_pack(w.medium, kernel.wavelength))
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
