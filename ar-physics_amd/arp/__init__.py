"""MI355X particle step for ar-physics scenes (SPH fluid, contact, elastic/plastic solids,
fused explicit RK stages) behind the C ABI of include/arp.h."""
from . import scene, tableau, wire  # noqa: F401
from ._lib import ArpError  # noqa: F401


def FluidSim(*args, **kw):
    from .sim import FluidSim as _F

    return _F(*args, **kw)
