"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product path never imports it.

States are dicts of numpy arrays {den (n), pos (n,4), vel (n,4), ref (ns,4),
stress (ns,16)} in float32 ("f") or float64 ("d").
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# (ARP_ORACLE_LIB: another build of the same sources, e.g. the sanitized one of `make sanitize`)
LIB = os.environ.get("ARP_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
VERLET = -1  # the stages field of a tableau that selects velocity Verlet

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)


class _AABB(C.Structure):
    _fields_ = [("min", C.c_float * 4), ("dim", C.c_float * 4)]


class _Scene(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("ns", C.c_uint32), ("nb", C.c_uint32), ("nm", C.c_uint32),
        ("mat", C.c_void_p), ("sid", C.c_void_p), ("part", C.c_void_p),
        ("bden", C.c_void_p), ("bpos", C.c_void_p), ("bvel", C.c_void_p), ("bmat", C.c_void_p),
        ("mats", C.c_void_p), ("inter", C.c_void_p),
        ("bounds", _AABB), ("h", C.c_float), ("g", C.c_float), ("threads", C.c_int),
        ("kscale", C.c_void_p),
    ]


class _State(C.Structure):
    _fields_ = [("den", C.c_void_p), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("ref", C.c_void_p), ("stress", C.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for nm in ("orc_compute_forces_f", "orc_compute_forces_d", "orc_fluid_forces_f",
                   "orc_fluid_forces_d", "orc_solid_forces_f", "orc_solid_forces_d"):
            getattr(L, nm).argtypes = [C.POINTER(_Scene), C.POINTER(_State), C.POINTER(_State)]
            getattr(L, nm).restype = C.c_int
        L.orc_strain_f.argtypes = [C.POINTER(_Scene), C.POINTER(_State), C.c_void_p]
        L.orc_strain_d.argtypes = [C.POINTER(_Scene), C.POINTER(_State), C.c_void_p]
        L.orc_step_f.argtypes = [C.POINTER(_Scene), C.c_void_p, C.c_float, C.c_int, C.POINTER(_State)]
        L.orc_step_d.argtypes = [C.POINTER(_Scene), C.c_void_p, C.c_double, C.c_int, C.POINTER(_State)]
        L.orc_lincomb_f.argtypes = [C.POINTER(_Scene), C.c_int, C.c_void_p, C.c_void_p, C.POINTER(_State)]
        L.orc_velocity_f.argtypes = [C.POINTER(_Scene), C.c_float, C.POINTER(_State), C.POINTER(_State)]
        L.orc_grid_dump.argtypes = [C.POINTER(_Scene), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_norm_const_f.argtypes = [C.c_uint32, C.c_float]
        L.orc_norm_const_f.restype = C.c_float
        L.orc_norm_const_d.argtypes = [C.c_uint32, C.c_double]
        L.orc_norm_const_d.restype = C.c_double
        L.orc_pressure_f.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float]
        L.orc_pressure_f.restype = C.c_float
        L.orc_contact_potential_f.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float]
        L.orc_contact_potential_f.restype = C.c_float
        L.orc_kernel_f.argtypes = [C.c_void_p, C.c_float, C.c_float]
        L.orc_kernel_f.restype = C.c_float
        L.orc_grad_w_f.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
        L.orc_rk_scalar.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_rk_scalar.restype = C.c_double
        L.orc_grid_sub.argtypes = [C.c_void_p, C.c_float, C.c_void_p]
        L.orc_dim.argtypes = [C.c_void_p]
        L.orc_dim.restype = C.c_uint32
        for nm in ("orc_seth_hill_f",):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_seth_hill_rate_f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_inverse_f.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_hooke_f.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p]
        L.orc_cell_coords.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


LIB_DUAL = os.path.join(HERE, "liboracle_dual.so")
_lib_dual = None
# parameter codes of the derivative build (the order of the adjoint's parameter gradients)
PARAM_FIELDS = {"bulk_visc": 1, "visc": 2, "normal_stiffness": 3, "shear_stiffness": 4,
                "normal_damp": 5, "shear_damp": 6}


def lib_dual():
    """The forward-mode derivative build (oracle_dual.cpp: REAL = an FP64 dual number)."""
    global _lib_dual
    if _lib_dual is None:
        if not os.path.exists(LIB_DUAL):
            build()
        L = C.CDLL(LIB_DUAL)
        L.orc_jvp_step.argtypes = [C.POINTER(_Scene), C.c_void_p, C.c_double, C.c_int, C.POINTER(_State),
                                   C.POINTER(_State), C.c_int, C.c_int]
        L.orc_jvp_step.restype = C.c_int
        _lib_dual = L
    return _lib_dual


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The oracle bound to one scene (the static part of FluidSim::new's state)."""

    def __init__(self, scene, threads=None, kscale=None):
        self.scene = scene
        p, b = scene.particles, scene.boundary
        self._keep = dict(
            mat=np.ascontiguousarray(p["mat"], np.uint32),
            sid=np.ascontiguousarray(p["solid_id"], np.uint32),
            part=np.ascontiguousarray(scene.solids["part_id"], np.uint32),
            bden=np.ascontiguousarray(b["den"], np.float32),
            bpos=np.ascontiguousarray(b["pos"], np.float32),
            bvel=np.ascontiguousarray(b["vel"], np.float32),
            bmat=np.ascontiguousarray(b["mat"], np.uint32),
            mats=np.ascontiguousarray(scene.materials),
            inter=np.ascontiguousarray(scene.interactions),
        )
        k = self._keep
        sc = _Scene()
        sc.n, sc.ns, sc.nb, sc.nm = len(p), len(scene.solids), len(b), len(scene.materials)
        for f in ("mat", "sid", "part", "bden", "bpos", "bvel", "bmat", "mats", "inter"):
            setattr(sc, f, k[f].ctypes.data)
        sc.bounds.min[:] = [float(x) for x in scene.bounds["min"]]
        sc.bounds.dim[:] = [float(x) for x in scene.bounds["dim"]]
        sc.h = scene.h
        sc.g = scene.g
        sc.threads = threads if threads is not None else (os.cpu_count() or 1)
        # optional per-solid stiffness scale (extension for gradient checks; None = reference)
        if kscale is not None:
            self._keep["kscale"] = np.ascontiguousarray(kscale, np.float32)
            sc.kscale = self._keep["kscale"].ctypes.data
        self.sc = sc

    # -- states -------------------------------------------------------------------------
    def state(self, prec="f", particles=None, solids=None):
        s = self.scene
        p = s.particles if particles is None else particles
        so = s.solids if solids is None else solids
        dt = np.float32 if prec == "f" else np.float64
        return dict(
            den=np.ascontiguousarray(p["den"], dt).copy(),
            pos=np.ascontiguousarray(p["pos"], dt).copy(),
            vel=np.ascontiguousarray(p["vel"], dt).copy(),
            ref=np.ascontiguousarray(so["ref_pos"], dt).copy(),
            stress=np.ascontiguousarray(so["stress"].reshape(len(so), 16), dt).copy(),
        )

    def zeros_like(self, st):
        return {k: np.zeros_like(v) for k, v in st.items()}

    @staticmethod
    def _cs(st):
        cs = _State()
        for k in ("den", "pos", "vel", "ref", "stress"):
            assert st[k].flags.c_contiguous
            setattr(cs, k, st[k].ctypes.data)
        return cs

    def _prec(self, st):
        return "f" if st["den"].dtype == np.float32 else "d"

    # -- calls --------------------------------------------------------------------------
    def compute_forces(self, st, which="all"):
        d = self.zeros_like(st)
        fn = {"all": "orc_compute_forces", "fluid": "orc_fluid_forces", "solid": "orc_solid_forces"}[which]
        rc = getattr(lib(), fn + "_" + self._prec(st))(C.byref(self.sc), C.byref(self._cs(st)), C.byref(self._cs(d)))
        assert rc == 0, rc
        return d

    def strain(self, st):
        out = np.zeros((len(st["ref"]), 3, 4, 4), st["den"].dtype)  # [s, k, col, row]
        fn = lib().orc_strain_f if self._prec(st) == "f" else lib().orc_strain_d
        rc = fn(C.byref(self.sc), C.byref(self._cs(st)), _p(out))
        assert rc == 0
        return out

    def step(self, st, tab, dt, nsteps=1):
        if int(tab["stages"]) == VERLET:
            return self.verlet(st, dt, nsteps)
        st = {k: v.copy() for k, v in st.items()}
        if self._prec(st) == "f":
            rc = lib().orc_step_f(C.byref(self.sc), _p(tab), C.c_float(dt), nsteps, C.byref(self._cs(st)))
        else:
            rc = lib().orc_step_d(C.byref(self.sc), _p(tab), C.c_double(dt), nsteps, C.byref(self._cs(st)))
        assert rc == 0, rc
        return st

    def jvp(self, st, tangent, tab, dt, nsteps=1, param=None):
        """Exact forward-mode derivative of nsteps steps (FP64): returns (y, ydot) with y the
        state after the steps from st and ydot = J tangent, plus dy/dp when param =
        (material index, field name of PARAM_FIELDS) seeds that parameter with tangent 1.
        st and tangent are FP64 states ("d")."""
        y = {k: np.ascontiguousarray(v, np.float64).copy() for k, v in st.items()}
        t = {k: np.ascontiguousarray(v, np.float64).copy() for k, v in tangent.items()}
        pm, pf = (-1, 0) if param is None else (int(param[0]), PARAM_FIELDS[param[1]])
        rc = lib_dual().orc_jvp_step(C.byref(self.sc), _p(tab), C.c_double(dt), int(nsteps), C.byref(self._cs(y)),
                                     C.byref(self._cs(t)), pm, pf)
        assert rc == 0, rc
        return y, t

    def verlet(self, st, dt, nsteps=1, D=None):
        """Velocity Verlet over this oracle's force function, in the state's precision: the
        integrator the reference names "verlet" (examples/test.rs:167) and steps with
        step_with_vel(t, state, dt, vel_fn, force_fn) (soft_body/mod.rs:252-265) after
        init_with_vel evaluated the forces once (soft_body/mod.rs:220-232).  The integrator
        crate (numerical-integration 0.0.1) is not vendored, so this is the textbook scheme
        x' = x + dt v + dt^2/2 a, v' = v + dt/2 (a + a'), with the same trapezoid for the
        density and the plastic strain, one force evaluation per step (parity unpinned
        beyond the properties tests/test_oracle.py checks).  The forces depend on velocity
        and density too, so a' is evaluated at the predicted v + dt a, rho + dt rho'
        (second order; the evaluation is kept for the next step):
            Z = (x + dt (v + dt/2 a), v + dt/2 a, rho + dt/2 rho', eps + dt/2 eps')
            X = (Z.x, Z.v + dt/2 a, Z.rho + dt/2 rho', Z.eps + dt/2 eps')
            a', rho'', eps'' = force_fn(X)
            y' = (Z.x, Z.v + dt/2 a', Z.rho + dt/2 rho'', Z.eps + dt/2 eps'')
        """
        s = {k: v.copy() for k, v in st.items()}
        f = s["den"].dtype.type
        dt = f(dt)
        h = f(0.5) * dt
        if D is None:
            D = self.compute_forces(s)
        for _ in range(nsteps):
            s["vel"][:, :3] = s["vel"][:, :3] + h * D["vel"][:, :3]
            s["pos"][:, :3] = s["pos"][:, :3] + dt * s["vel"][:, :3]
            s["den"] = s["den"] + h * D["den"]
            s["stress"] = s["stress"] + h * D["stress"]
            X = {k: v.copy() for k, v in s.items()}
            X["vel"][:, :3] = s["vel"][:, :3] + h * D["vel"][:, :3]
            X["den"] = s["den"] + h * D["den"]
            X["stress"] = s["stress"] + h * D["stress"]
            D = self.compute_forces(X)
            s["vel"][:, :3] = s["vel"][:, :3] + h * D["vel"][:, :3]
            s["den"] = s["den"] + h * D["den"]
            s["stress"] = s["stress"] + h * D["stress"]
        return s

    def lincomb(self, coefs, terms):
        out = self.zeros_like(terms[0])
        cs = [self._cs(t) for t in terms]
        arr = (C.POINTER(_State) * len(cs))(*[C.pointer(c) for c in cs])
        co = np.asarray(coefs, np.float32)
        rc = lib().orc_lincomb_f(C.byref(self.sc), len(terms), _p(co), C.cast(arr, C.c_void_p), C.byref(self._cs(out)))
        assert rc == 0
        return out

    def velocity(self, r, st):
        out = self.zeros_like(st)
        lib().orc_velocity_f(C.byref(self.sc), C.c_float(r), C.byref(self._cs(st)), C.byref(self._cs(out)))
        return out

    def grid_dump(self, pos=None, ref=None, with_counts=False, with_cand=True):
        s = self.scene
        n = s.n
        pos = np.ascontiguousarray(s.particles["pos"] if pos is None else pos, np.float32)
        ref = np.ascontiguousarray(s.solids["ref_pos"] if ref is None else ref, np.float32)
        cell = np.zeros(n, np.uint32)
        ref_cell = np.zeros(n, np.uint32)
        sub = grid_sub(s.bounds, s.h)
        ncells = int(sub[0]) * int(sub[1]) * int(sub[2])
        counts = np.zeros(ncells * 3, np.uint32) if with_counts else None
        cc = np.zeros(n, np.uint32)
        rc = lib().orc_grid_dump(C.byref(self.sc), _p(pos), _p(ref), _p(cell), _p(ref_cell),
                                 _p(counts) if counts is not None else None, _p(cc), None, 0)
        assert rc == 0
        cand = None
        if with_cand:
            tot = int(cc.sum(dtype=np.uint64))
            cand = np.zeros(max(tot, 1), np.uint32)
            rc = lib().orc_grid_dump(C.byref(self.sc), _p(pos), _p(ref), None, None, None, _p(cc),
                                     _p(cand), tot)
            assert rc == 0
            cand = cand[:tot]
        return dict(cell=cell, ref_cell=ref_cell, counts=counts, cand_count=cc, cand=cand)


def grid_sub(bounds, h):
    b = _AABB()
    b.min[:] = [float(x) for x in bounds["min"]]
    b.dim[:] = [float(x) for x in bounds["dim"]]
    sub = np.zeros(4, np.uint32)
    lib().orc_grid_sub(C.byref(b), C.c_float(h), _p(sub))
    return sub


def norm_const(dim, h, prec="f"):
    return lib().orc_norm_const_f(dim, h) if prec == "f" else lib().orc_norm_const_d(dim, h)


def rk_scalar(tab, lam, dt, y0):
    return lib().orc_rk_scalar(_p(tab), lam, dt, y0)


def rel_err(a, b):
    """Normwise relative error max|a-b| / max|b| (0 when both vanish)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)
