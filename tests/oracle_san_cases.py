"""Cases for the sanitized oracle build (run by tests/test_oracle_sanitized.py in a child
process with libasan preloaded and ARP_ORACLE_LIB pointing at oracle/_san/liboracle.so).
Every path of the oracle's C code runs once on small inputs: the grid with its edge cases
(NaN/inf, out of bounds, exactly at bounds.max, a bucket overflow), fluid forces in 2-D and
3-D with every equation of state and contact potential, solid forces with every strain order
and plastic flow, self contact, every integrator (velocity Verlet included), lin-comb with
chunks, the velocity function, in FP32 and FP64.  Not collected by pytest (no test_ prefix)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(HERE, "..", "ar-physics_amd"), os.path.join(HERE, "..")):
    sys.path.insert(0, os.path.abspath(p))

import numpy as np  # noqa: E402

from arp import scene as S  # noqa: E402
from arp import tableau as T  # noqa: E402
from helpers import deformed_solid_scene, random_grid_scene, self_contact_scene  # noqa: E402
from oracle import oracle as O  # noqa: E402

assert O.LIB == os.environ["ARP_ORACLE_LIB"], O.LIB


def run(name, sc, integs=("rk2",), dt=1e-5, steps=2):
    o = O.Oracle(sc, threads=2)
    o.grid_dump(with_counts=True)
    for prec in ("f", "d"):
        st = o.state(prec)
        o.compute_forces(st)
        for integ in integs:
            if integ == "verlet":
                o.verlet(st, dt, steps)
            else:
                o.step(st, T.tableau(integ), dt, steps)
    st = o.state("f")
    o.lincomb([0.5] * 7, [st] * 7)
    o.velocity(0.5, st)
    print("ok", name, flush=True)


for dim in (2, 3):
    run("grid edge cases %d-D" % dim, random_grid_scene(3, dim), ("euler", "rk4"), 1e-4, 1)
run("dam break 2-D", S.dam_break_2d(), ("rk4", "verlet"), 1e-4, 2)
sc = S.small_fluid_3d(n_side=6, box=(0.06, 0.06, 0.06))
run("column 3-D", sc, ("rk2", "heun3", "verlet"), 5e-5, 2)
# every equation of state and contact potential on the column
for eq in (0, 1, 2, 3):
    for pot in (0, 1, 2, 3, 4, 5):
        s2 = S.small_fluid_3d(n_side=5, box=(0.05, 0.05, 0.05))
        s2.materials["state_eq"][1] = eq
        s2.interactions["potential"][:] = pot
        s2.interactions["strength"][:] = 10.0
        s2.interactions["radius"][:] = s2.h
        s2.interactions["friction"][:] = 0.3
        run("eos %d contact %d" % (eq, pot), s2, ("rk2",), 5e-5, 1)
for order in (0, 1, 2, 3, -2):
    run("solid order %d" % order, deformed_solid_scene(3, order=order), ("rk4",), 1e-6, 1)
run("plastic solid", deformed_solid_scene(3, plastic=True), ("rk2", "verlet"), 1e-6, 2)
run("solid 2-D", deformed_solid_scene(2), ("rk2",), 1e-6, 1)
run("self contact", self_contact_scene(), ("rk2",), 1e-6, 1)
run("self contact outside the grid", self_contact_scene(outside=True), ("rk2",), 1e-6, 1)
print("all cases ok", flush=True)
