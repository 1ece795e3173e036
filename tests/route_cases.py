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
