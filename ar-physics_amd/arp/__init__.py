"""MI355X particle step for ar-physics scenes (SPH fluid, contact, elastic/plastic solids,
fused explicit RK stages) behind the C ABI of include/arp.h."""
from . import wire, scene, tableau  # noqa: F401
