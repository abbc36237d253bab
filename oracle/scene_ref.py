"""Scene-file oracle: a numpy restatement of the reference example runner's TOML scene
reader and of FluidSim::new's particle generation.  TEST INFRASTRUCTURE ONLY: imported by
tests/ to check the native scene-file reader against; the product path never imports it.

Restated from the reference (paths relative to the reference repository):

* Config::parse (examples/test_mods/config.rs:298-474): keys, defaults, the state-equation
  inference, the interaction tables and the material order (boundaries, then objects, each
  in byte order of the names, as toml 0.5's BTreeMap tables iterate);
* the regions of a table (config.rs:110-257) and their geometry (src/soft_body/
  material_region/region.rs:11-272): AABB, AABE, border shells, Transformed with the
  reference's Gauss-Jordan solve, Union / Difference / Intersection;
* STL model regions (examples/test_mods/mesh.rs:1-140): Mesh (crossing parity) and Surface;
* MaterialRegion::gen_particles (material_region/mod.rs:56-155): the staggered lattice in
  f32, the rigid velocity field (config.rs:405-418) and the lattice mass in f64;
* FluidSim::new (soft_body/mod.rs:58-111) and init_solid_particles (particles.rs:488-507).

The TOML text is read with ``tomli`` (a TOML 1.0 reader; the reference uses the toml 0.5
crate, not vendored).  Every f32 operation is a numpy float32 operation in the reference's
order, so the outputs are bit-exact restatements.  Pinned by tests/test_scene_file.py:
closed forms of every region, lattice spacings and offsets, masses of known boxes, and the
reference's own example scenes.
"""
import math
import os

import numpy as np

f32 = np.float32
INVALID = 0xFFFFFFFF


# ---------------------------------------------------------------------------------------
# regions: contains(p) takes an (n, 4) float32 array
# ---------------------------------------------------------------------------------------
class AABB:
    def __init__(self, mn, dim):
        self.min = np.zeros(4, f32)
        self.dim = np.zeros(4, f32)
        self.min[: len(mn)] = mn
        self.dim[: len(dim)] = dim

    def bound(self):
        return self

    def contains(self, p):
        # p[i] >= min[i] && p[i]-min[i] <= dim[i] on every axis (region.rs:64-71)
        ok = np.ones(len(p), bool)
        for i in range(4):
            ok &= ~(p[:, i] < self.min[i])
            ok &= ~((p[:, i] - self.min[i]) > self.dim[i])
        return ok

    def center(self):
        return np.array([self.min[i] + f32(0.5) * self.dim[i] for i in range(4)], f32)

    def max(self):
        return np.array([self.min[i] + self.dim[i] for i in range(4)], f32)

    def border(self, b):
        # AABB::border (region.rs:33-43)
        b = f32(b)
        c = self.center()
        inner_min = np.array([np.fmin(self.min[i] + b, c[i]) for i in range(4)], f32)
        inner_dim = np.array([np.fmax(self.dim[i] - f32(2.0) * b, f32(0.0)) for i in range(4)], f32)
        return Difference(self, AABB(inner_min, inner_dim))


def from_min_max(mn, mx):
    return AABB(mn, np.array([mx[i] - mn[i] for i in range(4)], f32))


class AABE:
    def __init__(self, center, radii):
        self.c = np.zeros(4, f32)
        self.r = np.zeros(4, f32)
        self.c[: len(center)] = center
        self.r[: len(radii)] = radii

    def bound(self):
        return AABB(self.c - self.r, f32(2.0) * self.r)

    def contains(self, p):
        # region.rs:88-106, evaluated in f64 like the reference
        lhs = np.zeros(len(p))
        for i in range(4):
            d = (p[:, i] - self.c[i]).astype(np.float64)
            r = float(self.r[i])
            if r > 0.0:
                coeff = 1.0
                for j in range(4):
                    if self.r[j] > 0.0 and j != i:
                        coeff *= float(self.r[j])
                lhs = lhs + d * d * coeff * coeff
        rhs = 1.0
        for i in range(4):
            if self.r[i] > 0.0:
                rhs *= float(self.r[i])
        return lhs <= rhs * rhs

    def border(self, b):
        inner = AABE(self.c, np.fmax(self.r - f32(b), f32(0)))
        return Difference(self, inner)


class Difference:
    def __init__(self, a, b):
        self.a, self.b = a, b

    def bound(self):
        return self.a.bound()

    def contains(self, p):
        return self.a.contains(p) & ~self.b.contains(p)


class Union:
    def __init__(self, a, b):
        self.a, self.b = a, b

    def bound(self):
        l, r = self.a.bound(), self.b.bound()
        mn = np.fmin(l.min, r.min)
        mx = np.fmax(l.max(), r.max())
        return from_min_max(mn, mx)

    def contains(self, p):
        return self.a.contains(p) | self.b.contains(p)


class Intersection:
    def __init__(self, a, b):
        self.a, self.b = a, b

    def bound(self):
        l, r = self.a.bound(), self.b.bound()
        mn = np.fmax(l.min, r.min)
        mx = np.fmin(l.max(), r.max())
        return AABB(mn, np.array([mx[i] - mn[i] for i in range(4)], f32))

    def contains(self, p):
        return self.a.contains(p) & self.b.contains(p)


class Transformed:
    """Transformed(region, m, t) (region.rs:126-225); m[j] is column j of the GLSL mat4, so
    y_i = sum_j m[j][i] x_j + t_i."""

    def __init__(self, base, m, t):
        self.base = base
        self.m = np.array(m, f32)
        self.t = np.zeros(4, f32)
        self.t[: len(t)] = t

    def bound(self):
        b = self.base.bound()
        m = self.m
        mn = self.t.copy()
        mx = self.t.copy()
        for i in range(4):
            a_min = f32(0.0)
            for j in range(4):
                a_min = f32(a_min + f32(m[j][i] * b.min[j]))
            mn[i] = f32(mn[i] + a_min)
            mx[i] = f32(mx[i] + a_min)
            for j in range(4):
                mn[i] = f32(mn[i] + np.fmin(f32(m[j][i] * b.dim[j]), f32(0.0)))
            for j in range(4):
                mx[i] = f32(mx[i] + np.fmax(f32(m[j][i] * b.dim[j]), f32(0.0)))
        return from_min_max(mn, mx)

    def contains(self, p):
        # the elimination's branches depend on the matrix only: replay them on all points
        a = self.m.copy()
        v = [(p[:, i] - self.t[i]).astype(f32) for i in range(4)]

        def row_mul(row, col, factor):
            for k in range(col, 4):
                a[k][row] = f32(a[k][row] * factor)
            v[row] = (v[row] * factor).astype(f32)

        def row_sum(row, col, factor, dest):
            for k in range(col, 4):
                a[k][dest] = f32(a[k][dest] + f32(factor * a[k][row]))
            v[dest] = (v[dest] + (factor * v[row]).astype(f32)).astype(f32)

        def swap_rows(r1, r2, col):
            for k in range(col, 4):
                a[k][r1], a[k][r2] = a[k][r2], a[k][r1]
            v[r1], v[r2] = v[r2], v[r1]

        i = j = 0
        while i < 4 and j < 4:
            if a[j][i] != 0.0:
                pivot = a[j][i]
                if pivot != 1.0:
                    row_mul(i, j, f32(f32(1.0) / pivot))
                for k in range(4):
                    if k != i and a[j][k] != 0.0:
                        row_sum(i, j, f32(-a[j][k]), k)
                i += 1
                j += 1
            else:
                cont = False
                for k in range(i + 1, 4):
                    if a[j][k] != 0.0:
                        swap_rows(i, k, j)
                        cont = True
                if not cont:
                    j += 1
        return self.base.contains(np.stack(v, 1).astype(f32))


def _signum(x):
    return np.where(np.isnan(x), x, np.copysign(f32(1.0), x)).astype(f32)


class TriMesh:
    def __init__(self, verts, faces, normals):
        self.verts = np.asarray(verts, f32).reshape(-1, 3)
        self.faces = np.asarray(faces, np.int64).reshape(-1, 3)
        self.normals = np.asarray(normals, f32).reshape(-1, 3)

    def bound(self):
        # point_bound (mesh.rs:13-24)
        mn = np.zeros(4, f32)
        mx = np.zeros(4, f32)
        mn[:3] = self.verts.min(0)
        mx[:3] = self.verts.max(0)
        return from_min_max(mn, mx)


class Mesh:
    """Mesh::contains (mesh.rs:27-72): parity of faces at or above the point whose xy
    projection holds it."""

    def __init__(self, mesh):
        self.mesh = mesh

    def bound(self):
        return self.mesh.bound()

    def contains(self, p):
        V = self.mesh.verts
        F = self.mesh.faces
        v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
        s1x, s1y = v1[:, 0] - v0[:, 0], v1[:, 1] - v0[:, 1]
        s2x, s2y = v2[:, 0] - v0[:, 0], v2[:, 1] - v0[:, 1]
        A = (s1x * s2y - s1y * s2x).astype(f32)
        sg = _signum(A)
        Aabs = np.abs(A)
        out = np.zeros(len(p), bool)
        for k in range(len(p)):
            x, y, z, w = p[k]
            if w != 0.0:
                continue
            above = (v0[:, 2] >= z) | (v1[:, 2] >= z) | (v2[:, 2] >= z)
            dx = (x - v0[:, 0]).astype(f32)
            dy = (y - v0[:, 1]).astype(f32)
            s = sg * (dx * s2y - dy * s2x)
            t = sg * (s1x * dy - s1y * dx)
            hit = above & (s >= 0) & (s <= Aabs) & (t >= 0) & (t <= Aabs) & ((s + t) <= Aabs)
            out[k] = bool(np.count_nonzero(hit) & 1)
        return out


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]).astype(f32)


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1).astype(f32)


class Surface:
    """Surface::contains (mesh.rs:88-137): within border of a vertex, or of a face inside
    its prism."""

    def __init__(self, mesh, border):
        self.mesh = mesh
        self.border = f32(border)

    def bound(self):
        b = self.mesh.bound()
        mn, dm = b.min.copy(), b.dim.copy()
        for i in range(3):
            mn[i] = f32(mn[i] - self.border)
            dm[i] = f32(dm[i] + f32(self.border * f32(2.0)))
        return AABB(mn, dm)

    def contains(self, p):
        V = self.mesh.verts
        F = self.mesh.faces
        N = self.mesh.normals
        a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
        bb = f32(self.border * self.border)
        s1, s2, s3, s4 = b - a, c - a, c - b, a - b
        n1, n2, n3 = _cross(s1, N), _cross(s2, N), _cross(s3, N)
        lim = (np.sqrt(_dot(N, N)) * self.border).astype(f32)
        e1, e2, e3 = _signum(_dot(s2, n1)), _signum(_dot(s1, n2)), _signum(_dot(s4, n3))
        out = np.zeros(len(p), bool)
        for k in range(len(p)):
            q = p[k, :3]
            d = (q - V).astype(f32)
            if np.any(_dot(d, d) <= bb):
                out[k] = True
                continue
            d = (q - a).astype(f32)
            near = np.abs(_dot(d, N)) <= lim
            d2 = (q - b).astype(f32)
            hit = near & (_signum(_dot(d, n1)) == e1) & (_signum(_dot(d, n2)) == e2) & (_signum(_dot(d2, n3)) == e3)
            out[k] = bool(np.any(hit))
        return out


def read_stl(path):
    """stl_io 0.4's read_stl: ASCII when the text opens with "solid" and the next line with
    "facet" or "endsolid", else binary (80-byte header, count, 50-byte records)."""
    data = open(path, "rb").read()
    ascii_ = False
    if data[:5] == b"solid":
        e = data.find(b"\n")
        if e >= 0:
            rest = data[e + 1:].lstrip()
            ascii_ = rest.startswith(b"facet") or rest.startswith(b"endsolid")
    tris, nrms = [], []
    if ascii_:
        words = data.split(b"\n", 1)[1].split()
        k = 0
        while k < len(words) and words[k] != b"endsolid":
            assert words[k] == b"facet" and words[k + 1] == b"normal"
            nrms.append([f32(float(x)) for x in words[k + 2:k + 5]])
            k += 7  # facet normal x y z outer loop
            vs = []
            for _ in range(3):
                assert words[k] == b"vertex"
                vs.append([f32(float(x)) for x in words[k + 1:k + 4]])
                k += 4
            k += 2  # endloop endfacet
            tris.append(vs)
        tris = np.array(tris, f32).reshape(-1, 9)
        nrms = np.array(nrms, f32).reshape(-1, 3)
    else:
        nt = int(np.frombuffer(data[80:84], "<u4")[0])
        rec = np.frombuffer(data[84:84 + 50 * nt], np.uint8).reshape(nt, 50)
        fl = np.ascontiguousarray(rec[:, :48]).view("<f4").reshape(nt, 12)
        nrms = fl[:, :3].astype(f32)
        tris = fl[:, 3:].astype(f32)
    # merge vertices by bit pattern, first appearance first
    flat = np.ascontiguousarray(tris.reshape(-1, 3))
    keys = flat.view(np.uint32).reshape(-1, 3)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    verts = flat[first[order]]
    faces = rank[inv.reshape(-1)].reshape(-1, 3)
    return TriMesh(verts, faces, nrms)


# ---------------------------------------------------------------------------------------
# the scene file (config.rs)
# ---------------------------------------------------------------------------------------
class ConfigError(ValueError):
    pass


def _float_or(t, k, default):
    if k not in t:
        return default
    v = t[k]
    if not isinstance(v, float):
        raise ConfigError("%s must be a float" % k)
    return v


def _int_or(t, k, default):
    if k not in t:
        return default
    v = t[k]
    if isinstance(v, bool) or not isinstance(v, int):
        raise ConfigError("%s must be an integer" % k)
    return v


def _vec4(a):
    out = np.zeros(4, f32)
    for i in range(min(4, len(a))):
        if not isinstance(a[i], float):
            raise ConfigError("vector entries must be floats")
        out[i] = f32(a[i])
    return out


def _vec4_or(t, k, default):
    return _vec4(t[k]) if k in t else np.array(default, f32)


_libm = None


def _cosf_sinf(x):
    """cosf / sinf of the C library, which Rust's f32::cos / f32::sin call on Linux."""
    global _libm
    if _libm is None:
        import ctypes

        _libm = ctypes.CDLL("libm.so.6")
        for nm in ("cosf", "sinf"):
            getattr(_libm, nm).argtypes = [ctypes.c_float]
            getattr(_libm, nm).restype = ctypes.c_float
    return f32(_libm.cosf(float(x))), f32(_libm.sinf(float(x)))


def _rotation_matrix(rot, scale):
    """parse_region's axis-angle matrix (config.rs:141-163), row 1's first entry as the
    reference writes it, row i scaled by scale[i]."""
    rad = f32(rot[3] * f32(f32(math.pi) / f32(180.0)))
    c, s = _cosf_sinf(rad)
    ul = np.sqrt(f32(f32(rot[0] * rot[0] + rot[1] * rot[1]) + rot[2] * rot[2]), dtype=f32)
    u = [f32(rot[0] / ul), f32(rot[1] / ul), f32(rot[2] / ul)]
    one_c = f32(f32(1.0) - c)
    m = [
        [c + u[0] * u[0] * one_c, u[0] * u[1] * one_c + u[2] * s, u[0] * u[2] * one_c - u[1] * s, f32(0.0)],
        [u[1] * u[2] * one_c - u[1] * s, c + u[1] * u[1] * one_c, u[1] * u[2] * one_c + u[0] * s, f32(0.0)],
        [u[0] * u[2] * one_c + u[1] * s, u[1] * u[2] * one_c - u[0] * s, c + u[2] * u[2] * one_c, f32(0.0)],
        [f32(0.0), f32(0.0), f32(0.0), f32(1.0)],
    ]
    m = np.array([[f32(x) for x in row] for row in m], f32)
    for i in range(4):
        for j in range(4):
            m[i][j] = f32(m[i][j] * scale[i])
    return m


class RegionReader:
    def __init__(self, base_dir=None):
        self.base_dir = base_dir
        self.models = {}

    def model(self, path):
        full = path if (not self.base_dir or path.startswith("/")) else os.path.join(self.base_dir, path)
        if full not in self.models:
            self.models[full] = read_stl(full)
        return self.models[full]

    def _fold(self, arr, cls):
        regs = [self.region(t) for t in arr]
        acc = regs[0]
        for r in regs[1:]:
            acc = cls(acc, r)
        return acc

    def from_mat(self, t):
        if isinstance(t.get("union"), list):
            return self._fold(t["union"], Union)
        if isinstance(t.get("difference"), list):
            return self._fold(t["difference"], Difference)
        if isinstance(t.get("intersection"), list):
            return self._fold(t["intersection"], Intersection)
        if isinstance(t.get("region"), dict):
            return self.region(t["region"])
        raise ConfigError("no region")

    def region(self, t):
        base = self.base(t)
        if not ("translation" in t or "scale" in t or "rotation" in t):
            return base
        trans = _vec4_or(t, "translation", [0, 0, 0, 0])
        r = t.get("rotation")
        rot = np.array([0, 0, 1, r], f32) if isinstance(r, float) else _vec4_or(t, "rotation", [0, 0, 1, 0])
        sc = t.get("scale")
        scale = np.ones(4, f32)
        if isinstance(sc, float):
            scale[:] = f32(sc)
        elif isinstance(sc, list):
            for i in range(min(4, len(sc))):
                scale[i] = f32(sc[i])
        return Transformed(base, _rotation_matrix(rot, scale), trans)

    def base(self, t):
        if isinstance(t.get("model"), str):
            if isinstance(t.get("border"), float):
                return Surface(self.model(t["model"]), f32(t["border"]))
            return Mesh(self.model(t["model"]))
        border = t.get("border")
        if isinstance(t.get("dim"), list):
            boxed, dim = True, _vec4(t["dim"])
        elif isinstance(t.get("radii"), list):
            boxed, dim = False, (_vec4(t["radii"]) * f32(2.0)).astype(f32)
        elif isinstance(t.get("diameters"), list):
            boxed, dim = False, _vec4(t["diameters"])
        else:
            return self.from_mat(t)
        mn = np.zeros(4, f32)
        if isinstance(t.get("min"), list):
            mn = _vec4(t["min"])
        elif isinstance(t.get("center"), list):
            mn = (_vec4(t["center"]) - (f32(0.5) * dim).astype(f32)).astype(f32)
        if boxed:
            box = AABB(mn, dim)
            return box.border(f32(border)) if border is not None else box
        dim = (dim * f32(0.5)).astype(f32)
        mn = (mn + dim).astype(f32)
        e = AABE(mn, dim)
        return e.border(f32(border)) if border is not None else e


POTENTIALS = {"Zero": 0, "zero": 0, "Constant": 1, "constant": 1, "Linear": 2, "linear": 2, "Tait": 3,
              "tait": 3, "Hertz": 5, "hertz": 5, "LennardJones": 4, "Lennard-Jones": 4, "Lennard Jones": 4,
              "lennard jones": 4, "lennard-jones": 4}


def _interaction(t, h):
    """parse_interation (config.rs:259-282); h is the f64 of the f32 kernel radius."""
    return dict(potential=POTENTIALS.get(t.get("potential"), 2),
                strength=f32(_float_or(t, "contact_strength", 10000.0)),
                radius=f32(_float_or(t, "contact_radius", _float_or(t, "contact_factor", 2.0) * h)),
                dampening=f32(_float_or(t, "dampening", 0.0)),
                friction=f32(_float_or(t, "friction", 0.0)))


MAT_FIELDS = ["immobile", "mass", "start_den", "state_eq", "sound_speed", "target_den", "visc", "bulk_visc",
              "strain_order", "normal_stiffness", "shear_stiffness", "normal_damp", "shear_damp", "plastic",
              "yield_strength", "work_hardening", "work_hardening_exp", "kinematic_hardening",
              "thermal_softening", "relaxation_time"]


def _material():
    m = {k: f32(0.0) for k in MAT_FIELDS}
    m["immobile"] = 0
    m["state_eq"] = 0
    m["strain_order"] = 0
    m["plastic"] = 0
    return m


def parse(doc, base_dir=None):
    """Config::parse of a tomli document: settings and the ordered object list."""
    rr = RegionReader(base_dir)
    cfg = dict(title=doc.get("name", "Fluid Test"), subticks=_int_or(doc, "subticks", 1),
               h=f32(_float_or(doc, "kernel_radius", 1.0 / 64.0)), dt=f32(_float_or(doc, "time_step", 0.01)),
               alpha=f32(_float_or(doc, "artificial_viscocity", 50.0)), g=f32(_float_or(doc, "gravity", 1.0)),
               min=_vec4_or(doc, "min", [-1, -1, 0, 0]), dim=_vec4_or(doc, "dim", [2, 2, 0, 0]),
               integrator=doc.get("integrator", "verlet"))
    h, alpha = cfg["h"], cfg["alpha"]
    objects, on_click, imap = [], [], {}

    def packing(t):
        return f32(_float_or(t, "packing", _float_or(t, "spacing", 0.5 * float(h)) / float(h)))

    def interactions(name, t):
        for k in sorted(t):
            if isinstance(t[k], dict) and k not in ("region", "union", "intersection", "difference"):
                imap[(name, k)] = _interaction(t[k], float(h))

    for name in sorted(doc.get("boundaries", {})):
        t = doc["boundaries"][name]
        m = _material()
        m.update(visc=f32(_float_or(t, "friction", 0.0)), start_den=f32(1.0), target_den=f32(1.0), immobile=1)
        objects.append(dict(name=name, packing=packing(t), region=rr.from_mat(t), mat=m, moving=False))
        interactions(name, t)
    for name in sorted(doc.get("objects", {})):
        t = doc["objects"][name]
        m = _material()
        den = _float_or(t, "density", 1.0)
        m["start_den"] = f32(_float_or(t, "start_density", den))
        m["target_den"] = f32(_float_or(t, "target_density", den))
        m["sound_speed"] = f32(_float_or(t, "speed_of_sound", 0.0))
        m["visc"] = f32(_float_or(t, "viscocity", _float_or(t, "friction", 0.0)))
        m["bulk_visc"] = f32(_float_or(t, "bulk_viscocity", float(f32(f32(alpha * h) * m["sound_speed"]))))
        st = t.get("state", t.get("state_equation"))
        if st is None:
            if "speed_of_sound" in t:
                m["state_eq"] = 2 if ("start_density" in t or "target_density" in t) else 3
        elif st in ("Tait", "tait", "Liquid", "liquid"):
            m["state_eq"] = 3
        elif st in ("Ideal_Gas", "ideal_gas", "gas", "Gas"):
            m["state_eq"] = 2
        elif st in ("Constant", "constant"):
            m["state_eq"] = 1
        else:
            raise ConfigError("Invalid state equation: %s" % st)
        m["strain_order"] = _int_or(t, "strain_order", 2)
        m["normal_stiffness"] = f32(_float_or(t, "normal_stiffness", 0.0))
        m["shear_stiffness"] = f32(_float_or(t, "shear_stiffness", 0.0))
        m["normal_damp"] = f32(_float_or(t, "normal_dampening", 0.0))
        m["shear_damp"] = f32(_float_or(t, "shear_dampening", 0.0))
        if "yield_strength" in t or "relaxation_time" in t:
            m["plastic"] = 1
            m["yield_strength"] = f32(_float_or(t, "yield_strength", 0.0))
            m["work_hardening"] = f32(_float_or(t, "work_hardening", 0.0))
            m["work_hardening_exp"] = f32(_float_or(t, "work_hardening_exp", 2.0))
            m["kinematic_hardening"] = f32(_float_or(t, "kinematic_hardening", 0.0))
            m["thermal_softening"] = f32(_float_or(t, "thermal_softening", 0.0))
            m["relaxation_time"] = f32(_float_or(t, "relaxation_time", 1.0))
        reg = rr.from_mat(t)
        ang = np.zeros(6, f32)
        for i, x in enumerate(t.get("angular_velocity", [])[:6]):
            ang[i] = f32(x)
        o = dict(name=name, packing=packing(t), region=reg, mat=m, moving=True,
                 vel=_vec4_or(t, "velocity", [0, 0, 0, 0]), ang=ang, center=reg.bound().center())
        if "on_click" in t:
            o["relative"] = t["on_click"] == "Relative"
            on_click.append(o)
        else:
            objects.append(o)
            interactions(name, t)
    no = len(objects)
    cfg_inter = {}
    for i in range(no):
        for j in range(no):
            a, b = objects[i]["name"], objects[j]["name"]
            it = imap.get((a, b), imap.get((b, a)))
            if it is not None:
                cfg_inter[(i, j)] = it
    cfg.update(objects=objects, on_click=on_click, inter=cfg_inter)
    return cfg


def lattice(bound, h):
    """The lattice points of gen_particles over a bounding box with spacing h (f32)."""
    mn, dm = bound.min, bound.dim
    h = f32(h)
    sq2, sq3, sq8 = np.sqrt(f32(2.0)), np.sqrt(f32(3.0)), np.sqrt(f32(8.0))
    rows = []
    pos = mn.copy()
    offset2 = f32(0.0)
    while True:
        offset = f32(0.0)
        pos[1] = f32(mn[1] + offset2)
        while True:
            x = f32(mn[0] + f32(offset + offset2))
            xs = [x]
            while True:
                x = f32(x + h)
                if f32(x - mn[0]) > dm[0] or dm[0] == f32(0.0):
                    break
                xs.append(x)
            row = np.empty((len(xs), 4), f32)
            row[:, 0] = xs
            row[:, 1] = pos[1]
            row[:, 2] = pos[2]
            row[:, 3] = pos[3]
            rows.append(row)
            pos[1] = f32(pos[1] + f32(h / sq2))
            offset = f32(h / f32(2.0)) if offset == f32(0.0) else f32(0.0)
            if f32(pos[1] - mn[1]) > dm[1] or dm[1] == f32(0.0):
                break
        pos[2] = f32(pos[2] + f32(h / sq3))
        offset2 = f32(h / sq8) if offset2 == f32(0.0) else f32(0.0)
        if f32(pos[2] - mn[2]) > dm[2] or dm[2] == f32(0.0):
            break
    return np.concatenate(rows)


def object_velocity(o, P):
    """The rigid velocity field of an [objects.*] table at points P (config.rs:405-418)."""
    if not o.get("moving"):
        return np.zeros_like(P)
    c, w, v = o["center"], o["ang"], o["vel"]
    r = (P - c).astype(f32)
    out = np.zeros_like(P)
    out[:, 0] = w[1] * r[:, 2] - w[2] * r[:, 1] - w[3] * r[:, 3] + v[0]
    out[:, 1] = w[2] * r[:, 0] - w[0] * r[:, 2] + w[4] * r[:, 3] + v[1]
    out[:, 2] = w[0] * r[:, 1] - w[1] * r[:, 0] - w[5] * r[:, 3] + v[2]
    out[:, 3] = w[3] * r[:, 0] - w[4] * r[:, 1] + w[5] * r[:, 3] + v[3]
    return out


def lattice_mass(start_den, bound, num):
    content = 1.0
    for i in range(4):
        if bound.dim[i] > 0.0:
            content = content * float(bound.dim[i])
    return f32((float(start_den) * content) / float(num))


def gen_object(o, h, mat_id):
    """MaterialRegion::gen_particles: (particles dict, mass, lattice size)."""
    sp = f32(o["packing"] * f32(h))
    b = o["region"].bound()
    L = lattice(b, sp)
    keep = o["region"].contains(L)
    P = L[keep]
    parts = dict(den=np.full(len(P), o["mat"]["start_den"], f32), mat=np.full(len(P), mat_id, np.uint32),
                 solid_id=np.full(len(P), INVALID, np.uint32), pos=P, vel=object_velocity(o, P))
    return parts, lattice_mass(o["mat"]["start_den"], b, len(L)), len(L)


def default_between(m1, m2, h):
    """MatInteraction::default_between (interaction.rs:42-50)."""
    if m1["state_eq"] != 0 and m2["state_eq"] != 0:
        return dict(potential=0, strength=f32(0), radius=f32(0), dampening=f32(0), friction=f32(0))
    return dict(potential=2, strength=f32(10000.0), radius=f32(f32(h) * f32(2.0)), dampening=f32(0),
                friction=f32(m1["visc"] + m2["visc"]))


def generate(cfg):
    """FluidSim::new's arrays: particles, boundary, solids, materials, interactions."""
    h = cfg["h"]
    mob, bnd, mats = [], [], []
    for i, o in enumerate(cfg["objects"]):
        parts, mass, _ = gen_object(o, h, i)
        m = dict(o["mat"])
        m["mass"] = mass
        mats.append(m)
        (bnd if m["immobile"] else mob).append(parts)

    def cat(lst):
        if not lst:
            return dict(den=np.zeros(0, f32), mat=np.zeros(0, np.uint32), solid_id=np.zeros(0, np.uint32),
                        pos=np.zeros((0, 4), f32), vel=np.zeros((0, 4), f32))
        return {k: np.concatenate([d[k] for d in lst]) for k in lst[0]}

    P, B = cat(mob), cat(bnd)
    stiff = np.array([m["normal_stiffness"] != 0 or m["shear_stiffness"] != 0 for m in mats], bool)
    idx = np.nonzero(stiff[P["mat"]])[0] if len(P["mat"]) else np.zeros(0, np.int64)
    P["solid_id"][idx] = np.arange(len(idx), dtype=np.uint32)
    if len(idx):
        S = dict(part_id=idx.astype(np.uint32), ref_pos=P["pos"][idx].copy())
    else:
        S = dict(part_id=np.array([INVALID], np.uint32), ref_pos=np.zeros((1, 4), f32))
    nm = len(mats)
    inter = [[None] * nm for _ in range(nm)]
    for i in range(nm):
        for j in range(i, nm):
            if (i, j) in cfg["inter"]:
                it = cfg["inter"][(i, j)]
            elif (j, i) in cfg["inter"]:
                it = cfg["inter"][(j, i)]
            elif i != j:
                it = default_between(mats[i], mats[j], h)
            else:
                it = dict(potential=0, strength=f32(0), radius=f32(0), dampening=f32(0), friction=f32(0))
            inter[i][j] = inter[j][i] = it
    return dict(particles=P, boundary=B, solids=S, materials=mats, inter=inter)


def add_particles(particles, solids, mats, inter, material, new, h):
    """Particles::add_particles (particle_state/particles.rs:547-631) on structured arrays in
    the wire layout: (particles, solids, materials, interactions, material id, added).

    A stiff material is always new; otherwise the first material equal field by field is
    reused.  A new material gets default_between with every material (particles.rs:573-598),
    with itself too unless it is stiff.  New particles of a stiff material become solids with
    their position as reference; a placeholder solid (part_id INVALID, the only one of a scene
    without solids) is replaced by them (the reference appends after it)."""
    nm = len(mats)
    solid = bool(material["normal_stiffness"] != 0 or material["shear_stiffness"] != 0)
    mid = nm
    if not solid:
        for i in range(nm):
            if all(mats[i][f] == material[f] for f in MAT_FIELDS):
                mid = i
                break
    added = mid == nm
    if added:
        mats = np.concatenate([mats, np.asarray(material, mats.dtype).reshape(1)])
        m2 = nm + 1
        J = np.zeros(m2 * m2, inter.dtype)

        def put(k, d):
            for f, v in d.items():
                J[k][f] = v

        for i in range(nm):
            for j in range(nm):
                J[i * m2 + j] = inter[i * nm + j]
            put(i * m2 + nm, default_between(mats[i], mats[nm], h))
        for i in range(nm):
            put(nm * m2 + i, default_between(mats[nm], mats[i], h))
        if not solid:
            put(nm * m2 + nm, default_between(mats[nm], mats[nm], h))
        inter = J
    new = np.array(new, copy=True)
    new["mat"] = mid
    new["solid_id"] = INVALID
    n = len(particles)
    if solid:
        if len(solids) == 1 and solids["part_id"][0] == INVALID:
            solids = solids[:0]
        extra = np.zeros(len(new), solids.dtype)
        extra["part_id"] = n + np.arange(len(new), dtype=np.uint32)
        extra["ref_pos"] = new["pos"]
        new["solid_id"] = len(solids) + np.arange(len(new), dtype=np.uint32)
        solids = np.concatenate([solids, extra])
    return np.concatenate([particles, new]), solids, mats, inter, mid, added
