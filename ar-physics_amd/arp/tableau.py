"""Explicit Runge-Kutta tableaus for the integrator names of examples/test.rs:165-180, and
the marker of velocity Verlet ("verlet", the TOML default, examples/test_mods/config.rs:451).

The reference takes them from the un-vendored ``numerical-integration 0.0.1`` crate; the
coefficients below are the textbook tableaus for those names.  "rk2" is taken to be the
explicit midpoint rule (the crate's own tableau for that name is not available), "heun2"
and "midpoint" are the two classic second-order methods, "rk3" is Kutta's third-order
method and "rk_3_8" is the 3/8 rule.  The same table is compiled into the C library
(``arp_tableau_by_name``).
"""
import numpy as np

from . import wire

_T = {
    "euler": ([[0.0]], [1.0]),
    "rk1": ([[0.0]], [1.0]),
    "midpoint": ([[0, 0], [0.5, 0]], [0.0, 1.0]),
    "rk2": ([[0, 0], [0.5, 0]], [0.0, 1.0]),
    "heun2": ([[0, 0], [1.0, 0]], [0.5, 0.5]),
    "heun": ([[0, 0], [1.0, 0]], [0.5, 0.5]),
    "ralston": ([[0, 0], [2.0 / 3.0, 0]], [0.25, 0.75]),
    "rk3": ([[0, 0, 0], [0.5, 0, 0], [-1.0, 2.0, 0]], [1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0]),
    "heun3": ([[0, 0, 0], [1.0 / 3.0, 0, 0], [0, 2.0 / 3.0, 0]], [0.25, 0.0, 0.75]),
    "rk4": (
        [[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]],
        [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0],
    ),
    "rk_3_8": (
        [[0, 0, 0, 0], [1.0 / 3.0, 0, 0, 0], [-1.0 / 3.0, 1.0, 0, 0], [1.0, -1.0, 1.0, 0]],
        [0.125, 0.375, 0.375, 0.125],
    ),
}

NAMES = sorted(_T) + ["verlet"]
VERLET = -1  # ArpTableau.stages of velocity Verlet (ARP_VERLET)


def tableau(name):
    if name.lower() == "verlet":
        t = np.zeros((), wire.TABLEAU)
        t["stages"] = VERLET
        return t
    a, b = _T[name.lower()]
    t = np.zeros((), wire.TABLEAU)
    s = len(b)
    t["stages"] = s
    for i in range(s):
        for j in range(s):
            t["a"][i, j] = a[i][j]
        t["b"][i] = b[i]
        t["c"][i] = sum(a[i])
    return t
