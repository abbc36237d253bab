"""The oracle's forward-mode derivative build (oracle/oracle_dual.cpp) against central
differences of the FP64 oracle: directional derivatives of three RK2 steps of a stirred
fluid and of a deformed solid next to a fluid, and a material parameter's derivative.  The
dual arithmetic is exact, so the only error is the differences' truncation and rounding
(steps of 1e-4 of the scale of each field, 0.1 kg/m^3 for the density)."""
import dataclasses

import numpy as np

from arp import scene as S
from arp import tableau as T
from helpers import deformed_solid_scene


def _fd_check(O, sc, seed, param=None):
    o = O.Oracle(sc)
    st = o.state("d")
    rng = np.random.default_rng(seed)
    st["vel"][:, :3] += 0.05 * rng.standard_normal((sc.n, 3))
    tab = T.tableau("rk2")
    dt, ns = 2e-5, 3
    worst = 0.0
    for field, scale in (("pos", 1e-3), ("vel", 1e-2), ("den", 1.0)):
        tan = {k: np.zeros_like(v) for k, v in st.items()}
        if field == "den":
            tan["den"][:] = rng.standard_normal(sc.n) * scale
        else:
            tan[field][:, :3] = rng.standard_normal((sc.n, 3)) * scale
        y, yd = o.jvp(st, tan, tab, dt, ns)
        y0 = o.step(st, tab, dt, ns)
        for k in ("den", "pos", "vel"):
            assert np.array_equal(y[k], y0[k])  # the values are the FP64 oracle's
        # (a density tangent moves positions only through the m_i/rho_i-scaled pressure: a
        # larger step keeps the differences above the positions' rounding)
        eps = 1e-1 if field == "den" else 1e-4
        yp = o.step({k: st[k] + eps * tan[k] for k in st}, tab, dt, ns)
        ym = o.step({k: st[k] - eps * tan[k] for k in st}, tab, dt, ns)
        for k in ("pos", "vel"):
            fd = (yp[k] - ym[k]) / (2 * eps)
            err = np.abs(fd - yd[k]).max() / np.abs(fd).max()
            worst = max(worst, err)
            assert err < 1e-5, (field, k, err)
    if param is not None:
        m, name = param
        _, yd = o.jvp(st, {k: np.zeros_like(v) for k, v in st.items()}, tab, dt, ns, param=param)
        base = float(sc.materials[name][m])
        dp = abs(base) * 1e-4
        res = []
        for sgn in (1, -1):
            mm = sc.materials.copy()
            mm[name][m] = base + sgn * dp
            res.append(O.Oracle(dataclasses.replace(sc, materials=mm)).step(st, tab, dt, ns))
        dpe = (float(base + dp) - float(base - dp))
        fd = (res[0]["vel"] - res[1]["vel"]) / dpe
        err = np.abs(fd - yd["vel"]).max() / np.abs(fd).max()
        assert err < 2e-3, (name, err)
    return worst


def test_dual_fluid(oracle_mod):
    sc = S.small_fluid_3d(n_side=6, box=(0.05, 0.05, 0.05))
    _fd_check(oracle_mod, sc, 0, param=(1, "bulk_visc"))


def test_dual_solid(oracle_mod):
    sc = deformed_solid_scene(dim=3, seed=3)
    _fd_check(oracle_mod, sc, 1, param=(2, "shear_stiffness"))
