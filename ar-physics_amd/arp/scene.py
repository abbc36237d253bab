"""Scene construction: the host-side derivations of ``FluidSim::new`` and the workloads.

Restated from the reference (paths relative to the reference repository):

* material constructors (material.rs:61-121, examples/test_mods/config.rs:330-391);
* the interaction matrix (soft_body/mod.rs:93-111, interaction.rs:42-50);
* the grid dimensionality (soft_body/mod.rs:113-117).

Regions and the staggered lattice are generated by the native scene-file reader
(``arp.config``); the 2-D dam break is a TOML scene read by it.  The 3-D benchmark workloads
(fluid column, elastic cantilever, fluid-solid contact, ensembles) are built here from their
stated sizes and parameters.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import wire

f32 = np.float32


# ---------------------------------------------------------------------------------------
# materials and interactions
# ---------------------------------------------------------------------------------------
def immobile_material(friction=0.0):
    """Material::new_immobile (material.rs:113-120)."""
    return wire.material(visc=friction, start_den=1.0, target_den=1.0, immobile=1)


def fluid_material(kernel_radius, alpha, density=1.0, speed_of_sound=None, viscosity=0.0,
                   start_density=None, target_density=None, state=None, bulk_viscosity=None):
    """An [objects.*] fluid table as config.rs:354-375 parses it (state inferred when not
    given: speed_of_sound alone -> Tait, with start/target density -> ideal gas)."""
    h = f32(kernel_radius)
    c = f32(speed_of_sound if speed_of_sound is not None else 0.0)
    m = wire.material()
    m["start_den"] = start_density if start_density is not None else density
    m["target_den"] = target_density if target_density is not None else density
    m["sound_speed"] = c
    m["visc"] = viscosity
    if bulk_viscosity is None:
        bulk_viscosity = float(f32(f32(f32(alpha) * h) * c))
    m["bulk_visc"] = bulk_viscosity
    if state is None:
        if speed_of_sound is not None:
            state = wire.EQ_IDEAL_GAS if (start_density is not None or target_density is not None) else wire.EQ_TAIT
        else:
            state = wire.EQ_ZERO
    m["state_eq"] = state
    m["strain_order"] = 2
    return m


def solid_material(density=1.0, normal_stiffness=0.0, shear_stiffness=0.0, normal_damp=0.0,
                   shear_damp=0.0, strain_order=2, plastic=None, viscosity=0.0):
    """An elastic/plastic [objects.*] table (config.rs:354-391)."""
    m = wire.material()
    m["start_den"] = density
    m["target_den"] = density
    m["visc"] = viscosity
    m["strain_order"] = strain_order
    m["normal_stiffness"] = normal_stiffness
    m["shear_stiffness"] = shear_stiffness
    m["normal_damp"] = normal_damp
    m["shear_damp"] = shear_damp
    if plastic is not None:
        m["plastic"] = 1
        m["yield_strength"] = plastic.get("yield_strength", 0.0)
        m["work_hardening"] = plastic.get("work_hardening", 0.0)
        m["work_hardening_exp"] = plastic.get("work_hardening_exp", 2.0)
        m["kinematic_hardening"] = plastic.get("kinematic_hardening", 0.0)
        m["thermal_softening"] = plastic.get("thermal_softening", 0.0)
        m["relaxation_time"] = plastic.get("relaxation_time", 1.0)
    return m


def lame_from_young(E, nu, sign=-1.0):
    """(lambda, mu) for Young's modulus E and Poisson ratio nu, with the reference's
    negative stiffness sign convention (all example solids use negative values)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return sign * lam, sign * mu


def default_interaction(m1, m2, h):
    """MatInteraction::default_between (interaction.rs:42-50)."""
    if m1["state_eq"] != wire.EQ_ZERO and m2["state_eq"] != wire.EQ_ZERO:
        return wire.interaction()
    return wire.interaction(wire.POT_LINEAR, 10000.0, float(f32(h) * f32(2.0)), 0.0,
                            float(f32(m1["visc"]) + f32(m2["visc"])))


def resolve_interactions(mats, h, configured=None):
    """FluidSim::new's interaction matrix (soft_body/mod.rs:93-111).

    configured: dict {(i, j): interaction} of explicitly configured pairs."""
    configured = configured or {}
    nm = len(mats)
    inter = np.zeros(nm * nm, wire.INTERACTION)
    for i in range(nm):
        for j in range(i, nm):
            if (i, j) in configured:
                it = configured[(i, j)]
            elif (j, i) in configured:
                it = configured[(j, i)]
            elif i != j:
                it = default_interaction(mats[i], mats[j], h)
            else:
                it = wire.interaction()
            inter[i * nm + j] = it
            inter[j * nm + i] = it
    return inter


def grid_dim(bounds):
    """dim = number of axes with bounds.dim > 0 (soft_body/mod.rs:113-117)."""
    return int(sum(1 for i in range(4) if bounds["dim"][i] > 0))


# ---------------------------------------------------------------------------------------
# scene container
# ---------------------------------------------------------------------------------------
@dataclasses.dataclass
class Scene:
    particles: np.ndarray  # wire.PARTICLE, mobile
    boundary: np.ndarray  # wire.PARTICLE, immobile
    materials: np.ndarray  # wire.MATERIAL
    interactions: np.ndarray  # wire.INTERACTION, nm*nm
    bounds: np.ndarray  # wire.AABB
    h: float
    g: float
    alpha: float = 0.0
    solids: np.ndarray = None  # wire.SOLID
    dt: float = 1e-4
    subticks: int = 1
    integrator: str = "rk4"
    name: str = ""
    objects: list = None  # per-object records of a scene file (wire.SCENE_OBJECT)

    def __post_init__(self):
        if self.solids is None:
            self.solids = init_solids(self.particles, self.materials)

    @property
    def n(self):
        return len(self.particles)

    @property
    def nb(self):
        return len(self.boundary)

    @property
    def ns(self):
        return len(self.solids)

    @property
    def dim(self):
        return grid_dim(self.bounds)

    def copy(self):
        return dataclasses.replace(
            self,
            particles=self.particles.copy(),
            boundary=self.boundary.copy(),
            solids=self.solids.copy(),
            materials=self.materials.copy(),
            interactions=self.interactions.copy(),
            bounds=self.bounds.copy(),
        )


def init_solids(particles, materials):
    """Particles::init_solid_particles (particles.rs:488-506): a SolidParticle for every
    particle of a stiff material, in particle order; one dummy when there is none."""
    stiff = (materials["normal_stiffness"] != 0) | (materials["shear_stiffness"] != 0)
    is_solid = stiff[particles["mat"]] if len(particles) else np.zeros(0, bool)
    idx = np.nonzero(is_solid)[0]
    if len(idx) == 0:
        s = wire.solids(1)
        s["part_id"] = wire.INVALID
        return s
    s = wire.solids(len(idx))
    s["part_id"] = idx.astype(np.uint32)
    s["ref_pos"] = particles["pos"][idx]
    particles["solid_id"][:] = wire.INVALID
    particles["solid_id"][idx] = np.arange(len(idx), dtype=np.uint32)
    return s


# ---------------------------------------------------------------------------------------
# deterministic RNG (splitmix64 -> uniform [-1, 1))
# ---------------------------------------------------------------------------------------
def splitmix64(seed, n):
    M = (1 << 64) - 1
    state = np.uint64(seed)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = state + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    del M
    return z


def uniform_pm1(seed, n):
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))) * 2.0 - 1.0


# ---------------------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------------------
DAM_BREAK_2D = """\
name = "Dam break 2D"
kernel_radius = {h!r}
time_step = 0.0001
gravity = 9.81
artificial_viscocity = {alpha!r}
integrator = "RK4"
min = [-0.03, -0.03]
dim = [1.06, 1.06]

[boundaries.wall]
spacing = {s!r}
region = {{min = [-0.03, -0.03], dim = [1.06, 1.06], border = {border!r}}}

[objects.water]
density = 1000.0
speed_of_sound = 20.0
viscocity = 0.0
spacing = {s!r}
region = {{min = [0.0, 0.0], dim = [{wx!r}, {wy!r}]}}
"""


def dam_break_2d(spacing=0.01, kernel_radius=0.013, rows=64, cols=64, alpha=0.02):
    """2-D weakly compressible dam break, read from a TOML scene (DAM_BREAK_2D) by the
    native scene-file reader: a wall shell of three lattice rows around [-0.03, 1.03]^2 and
    a water block of cols x rows lattice points (spacing = lattice step) at the origin.
    Materials follow the loader's order: boundaries, then objects, each sorted by name."""
    from . import config

    s = float(spacing)
    text = DAM_BREAK_2D.format(h=float(kernel_radius), alpha=float(alpha), s=s, border=3 * s,
                               wx=(cols - 0.25) * s, wy=(rows - 0.5) * s / math.sqrt(2.0))
    sc = config.loads(text)
    sc.integrator = "rk4"
    sc.name = "dam_break_2d"
    return sc


def _cubic_block(n, spacing, origin, jitter, seed):
    i = np.arange(n[0])
    j = np.arange(n[1])
    k = np.arange(n[2])
    I, J, K = np.meshgrid(i, j, k, indexing="ij")
    P = np.stack([I.ravel(), J.ravel(), K.ravel()], 1).astype(np.float64)
    P = (P + 0.5) * spacing + np.asarray(origin)
    if jitter:
        u = uniform_pm1(seed, P.size).reshape(P.shape)
        P = P + jitter * spacing * u
    return P


def _walls(box, spacing, layers, open_top=True):
    """Boundary lattice: cell centres of the box grown by `layers` that lie outside it."""
    lo = -layers
    hi = [int(round(box[a] / spacing)) + layers for a in range(3)]
    if open_top:
        hi[1] = int(round(box[1] / spacing))
    ax = [np.arange(lo, hi[a]) for a in range(3)]
    I, J, K = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    P = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    inner = [int(round(box[a] / spacing)) for a in range(3)]
    inside = np.ones(len(P), bool)
    for a in range(3):
        inside &= (P[:, a] >= 0) & (P[:, a] < inner[a])
    P = P[~inside]
    return (P.astype(np.float64) + 0.5) * spacing


def fluid_column_3d(shape=(128, 128, 64), spacing=0.005, jitter=0.05, seed=0, box=(1.5, 1.0, 1.5),
                    layers=3, alpha=0.02, h_factor=1.3, rk="rk2", dt=5e-5, n_side=None):
    """3-D fluid column collapse: a shape[0] x shape[1] x shape[2] (x, y up, z) cubic lattice
    of 1,048,576 particles by default, with seeded uniform jitter of +-jitter*spacing, in the
    corner of an open-top box with `layers` of fixed boundary particles; poly6 support
    h = h_factor*spacing; Tait (rho0 1000, gamma 7, c0 20); artificial viscosity alpha ->
    bulk_visc = alpha*h*c0.  n_side gives an n_side^3 block instead."""
    if n_side is not None:
        shape = (n_side,) * 3
    h = f32(h_factor * spacing)
    P = _cubic_block(tuple(shape), spacing, (0.0, 0.0, 0.0), jitter, seed)
    B = _walls(box, spacing, layers)
    water = fluid_material(h, alpha, density=1000.0, speed_of_sound=20.0, viscosity=0.0)
    wall = immobile_material(0.0)
    wall["mass"] = f32(1.0 * spacing**3)
    water["mass"] = f32(1000.0 * spacing**3)
    mats = np.array([wall, water], wire.MATERIAL)
    flu = wire.particles(len(P))
    flu["pos"][:, :3] = P
    flu["den"] = 1000.0
    flu["mat"] = 1
    bnd = wire.particles(len(B))
    bnd["pos"][:, :3] = B
    bnd["den"] = 1.0
    bnd["mat"] = 0
    margin = (layers + 1) * spacing
    bounds = wire.aabb([-margin] * 3, [box[0] + 2 * margin, box[1] + 2 * margin, box[2] + 2 * margin])
    inter = resolve_interactions(mats, h)
    return Scene(flu, bnd, mats, inter, bounds, float(h), 9.81, alpha, dt=dt, subticks=1,
                 integrator=rk, name="fluid_column_3d")


def fluid_block_3d(n, alpha=0.02, spacing=0.005, jitter=0.05, seed=0, layers=3, h_factor=1.3, rk="rk2",
                   dt=5e-5, name="fluid_block_3d"):
    """A 3-D block of exactly n fluid particles (the first n points of a jittered cubic
    lattice of about n^(1/3) per side, taken in lattice order) in the corner of an
    open-top box twice the block's width and 1.5 times its height, with `layers` of boundary
    particles; Tait (rho0 1000, c0 20), bulk_visc = alpha h c0."""
    side = int(np.ceil(n ** (1.0 / 3.0)))
    shape = (side, side, int(np.ceil(n / (side * side))))
    h = f32(h_factor * spacing)
    P = _cubic_block(shape, spacing, (0.0, 0.0, 0.0), jitter, seed)[:n]
    ext = [shape[a] * spacing for a in range(3)]
    box = (2.0 * ext[0], 1.5 * ext[1], 2.0 * ext[2])
    B = _walls(box, spacing, layers)
    water = fluid_material(h, alpha, density=1000.0, speed_of_sound=20.0, viscosity=0.0)
    wall = immobile_material(0.0)
    wall["mass"] = f32(1.0 * spacing**3)
    water["mass"] = f32(1000.0 * spacing**3)
    mats = np.array([wall, water], wire.MATERIAL)
    flu = wire.particles(len(P))
    flu["pos"][:, :3] = P
    flu["den"] = 1000.0
    flu["mat"] = 1
    bnd = wire.particles(len(B))
    bnd["pos"][:, :3] = B
    bnd["den"] = 1.0
    bnd["mat"] = 0
    margin = (layers + 1) * spacing
    bounds = wire.aabb([-margin] * 3, [box[0] + 2 * margin, box[1] + 2 * margin, box[2] + 2 * margin])
    inter = resolve_interactions(mats, h)
    return Scene(flu, bnd, mats, inter, bounds, float(h), 9.81, alpha, dt=dt, subticks=1, integrator=rk, name=name)


def ensemble_3d(count=256, nmin=2048, nmax=65536, alpha=(0.01, 0.1), seed=0, **kw):
    """Independent fluid scenes for batched runs: particle counts log-uniform in [nmin, nmax]
    and artificial viscosity alpha uniform in `alpha`, both drawn from the splitmix64 stream
    of `seed`; scene k uses lattice jitter seed seed + 1 + k."""
    u = (uniform_pm1(seed, 2 * count).reshape(count, 2) + 1.0) * 0.5  # [0, 1)
    ns = np.exp(np.log(nmin) + u[:, 0] * (np.log(nmax) - np.log(nmin)))
    ns = np.clip(np.floor(ns).astype(np.int64), nmin, nmax)
    al = alpha[0] + u[:, 1] * (alpha[1] - alpha[0])
    return [fluid_block_3d(int(n), alpha=float(a), seed=seed + 1 + k, name="ensemble_%d" % k, **kw)
            for k, (n, a) in enumerate(zip(ns, al))]


def small_fluid_3d(n_side=12, spacing=0.005, seed=1, jitter=0.05, box=(0.12, 0.1, 0.12),
                   rk="rk2", dt=5e-5):
    """A small 3-D column of the same parameters, for tests that run the oracle."""
    return fluid_column_3d(n_side=n_side, spacing=spacing, jitter=jitter, seed=seed, box=box,
                           rk=rk, dt=dt)


def cantilever_3d(nx=40, ny=40, nz=400, spacing=0.0025, E=1e6, nu=0.3, density=1000.0,
                  h_factor=1.3, clamp_layers=3, clamp_length=8, clamp_scale=3e6, dt=1e-5):
    """3-D elastic cantilever: an nx*ny*nz block of solid particles along z; linear elastic
    (St Venant-Kirchhoff on the Green-Lagrange strain, strain_order 2) with lambda, mu from E
    and nu in the reference's negative sign convention; gravity loading.  The end z < 0 is
    clamped by fixed boundary particles, `clamp_layers` thick: a wall behind the end face and
    plates above, below and beside the first `clamp_length` layers of the beam.  The
    reference has no clamp; the beam-clamp pair is given the Tait contact potential
    (fluid.rs:58), which repels closer than r0 and attracts beyond it, with r0 = spacing (the
    gap between the beam and the plates) and scale clamp_scale, so the plates grip the root
    from every side.  (fluid.rs:177 passes an interaction's radius as the potential's scale
    and its strength as r0.)  The default linear contact (interaction.rs:42-50) pushes with a
    nearly constant pressure and cannot hold the beam."""
    h = f32(h_factor * spacing)
    P = _cubic_block((nx, ny, nz), spacing, (0.0, 0.0, 0.0), 0.0, 0)
    lam, mu = lame_from_young(E, nu)
    solid = solid_material(density=density, normal_stiffness=lam, shear_stiffness=mu)
    solid["mass"] = f32(density * spacing**3)
    wall = immobile_material(0.0)
    wall["mass"] = f32(1.0 * spacing**3)
    mats = np.array([wall, solid], wire.MATERIAL)
    # the clamp: the lattice cells of the box around the root, grown by clamp_layers, minus
    # the beam itself
    L = clamp_layers
    ax = [np.arange(-L, nx + L), np.arange(-L, ny + L), np.arange(-L, clamp_length)]
    I, J, K = np.meshgrid(*ax, indexing="ij")
    Cg = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    beam = (Cg[:, 0] >= 0) & (Cg[:, 0] < nx) & (Cg[:, 1] >= 0) & (Cg[:, 1] < ny) & (Cg[:, 2] >= 0)
    C = (Cg[~beam].astype(np.float64) + 0.5) * spacing
    flu = wire.particles(len(P))
    flu["pos"][:, :3] = P
    flu["den"] = density
    flu["mat"] = 1
    bnd = wire.particles(len(C))
    bnd["pos"][:, :3] = C
    bnd["den"] = 1.0
    bnd["mat"] = 0
    margin = (clamp_layers + 2) * spacing
    ext = (nx * spacing, ny * spacing, nz * spacing)
    bounds = wire.aabb([-margin, -margin - 0.5 * ext[2], -margin],
                       [ext[0] + 2 * margin, ext[1] + 2 * margin + 0.5 * ext[2], ext[2] + 2 * margin])
    clamp = wire.interaction(wire.POT_TAIT, float(f32(spacing)), clamp_scale, 0.0, 0.0)
    inter = resolve_interactions(mats, h, {(0, 1): clamp})
    return Scene(flu, bnd, mats, inter, bounds, float(h), 9.81, 0.0, dt=dt, subticks=1,
                 integrator="rk4", name="cantilever_3d")


def supported_ball(n, spacing):
    """Exactly n points of a cubic lattice (spacing `spacing`) forming a ball, every one with at
    least one lattice neighbour along each axis, so that the corrected-gradient matrix
    L = sum V grad W(dX) dX^T of every solid particle (strain.rs:101-136) is invertible for a
    support radius h in (spacing, sqrt(2) spacing].  For radii from the ball of n points up
    in quarter spacings: the lattice ball with the unsupported points dropped (repeatedly, a
    drop can unsupport a neighbour) is trimmed to n by removing outermost points (by distance,
    then lattice index) whose removal leaves every neighbour supported; the first radius that
    reaches exactly n is taken.  Returns (n, 3) offsets from the centre."""
    r = (3.0 * n / (4.0 * math.pi)) ** (1.0 / 3.0)
    for _ in range(40):
        m = int(math.ceil(r)) + 1
        g1 = np.arange(-m, m + 1)
        I, J, K = np.meshgrid(g1, g1, g1, indexing="ij")
        G = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
        d2 = (G.astype(np.float64) ** 2).sum(1)
        keep = d2 <= r * r
        G, d2 = G[keep], d2[keep]
        order = np.lexsort((np.arange(len(G)), d2))
        G = G[order]
        idx = {tuple(g): k for k, g in enumerate(G)}
        nb = np.full((len(G), 6), -1, np.int64)  # neighbour along +x -x +y -y +z -z
        for k, g in enumerate(G):
            for q in range(6):
                h = [int(g[0]), int(g[1]), int(g[2])]
                h[q // 2] += 1 if q % 2 == 0 else -1
                nb[k, q] = idx.get(tuple(h), -1)
        inside = np.ones(len(G), bool)

        def has(k, q):
            j = nb[k, q]
            return j >= 0 and inside[j]

        def covered(k):
            return all(has(k, 2 * ax) or has(k, 2 * ax + 1) for ax in range(3))

        changed = True
        while changed:
            changed = False
            for k in np.nonzero(inside)[0]:
                if not covered(k):
                    inside[k] = False
                    changed = True
        count = int(inside.sum())
        progress = True
        while count > n and progress:
            progress = False
            for k in range(len(G) - 1, -1, -1):  # outermost first
                if not inside[k]:
                    continue
                # removable if every neighbour keeps a neighbour along the axis it shares
                # with k (k sits at the neighbour's side q ^ 1)
                if all(not (nb[k, q] >= 0 and inside[nb[k, q]]) or has(nb[k, q], q) for q in range(6)):
                    inside[k] = False
                    count -= 1
                    progress = True
                    break  # again from the outermost
        if count == n:
            assert all(covered(k) for k in np.nonzero(inside)[0])
            return G[inside].astype(np.float64) * spacing
        r += 0.25
    raise ValueError("no supported lattice ball of %d points" % n)


def contact_3d(n_fluid_side=64, n_sphere=65536, spacing=None, E=5e5, nu=0.4, density=1000.0,
               v0=(0.0, -2.0, 0.0), contact_strength=1e4, alpha=0.02, h_factor=1.3, dt=2e-5,
               seed=0, gap=1.0):
    """Fluid-solid contact: an n_fluid_side^3 jittered fluid block and an elastic ball of
    exactly n_sphere particles (cubic lattice, every particle with a full 3-D reference
    support, supported_ball) whose lowest layer is `gap` spacings above the fluid's top layer,
    moving at v0; linear contact potential of strength contact_strength (radius 2h, the
    loader's defaults).  Pairs interact within h only (the reference's 27-bucket sweep), so
    with gap = 1 < h / spacing the contact acts from the first step."""
    if spacing is None:
        spacing = 0.005
    h = f32(h_factor * spacing)
    Pf = _cubic_block((n_fluid_side,) * 3, spacing, (0.0, 0.0, 0.0), 0.05, seed)
    L = n_fluid_side * spacing
    Qo = supported_ball(n_sphere, spacing)
    r = float(np.abs(Qo).max())
    # the ball's lowest layer `gap` spacings above the fluid's top layer (at L - spacing / 2)
    c = np.array([0.5 * L, L + r + (gap - 0.5) * spacing, 0.5 * L])
    Q = Qo + c
    box = (L, L + 2 * r + 4 * spacing, L)
    B = _walls(box, spacing, 3)
    water = fluid_material(h, alpha, density=1000.0, speed_of_sound=20.0, viscosity=0.0)
    water["mass"] = f32(1000.0 * spacing**3)
    lam, mu = lame_from_young(E, nu)
    ball = solid_material(density=density, normal_stiffness=lam, shear_stiffness=mu)
    ball["mass"] = f32(density * spacing**3)
    wall = immobile_material(0.0)
    wall["mass"] = f32(spacing**3)
    mats = np.array([wall, ball, water], wire.MATERIAL)
    n = len(Pf) + len(Q)
    flu = wire.particles(n)
    flu["pos"][: len(Pf), :3] = Pf
    flu["den"][: len(Pf)] = 1000.0
    flu["mat"][: len(Pf)] = 2
    flu["pos"][len(Pf):, :3] = Q
    flu["vel"][len(Pf):, :3] = v0
    flu["den"][len(Pf):] = density
    flu["mat"][len(Pf):] = 1
    bnd = wire.particles(len(B))
    bnd["pos"][:, :3] = B
    bnd["den"] = 1.0
    bnd["mat"] = 0
    configured = {(1, 2): wire.interaction(wire.POT_LINEAR, contact_strength, float(f32(2.0) * h), 0.0, 0.0)}
    inter = resolve_interactions(mats, h, configured)
    margin = 4 * spacing
    bounds = wire.aabb([-margin] * 3, [box[0] + 2 * margin, box[1] + 2 * margin, box[2] + 2 * margin])
    return Scene(flu, bnd, mats, inter, bounds, float(h), 9.81, alpha, dt=dt, subticks=1,
                 integrator="rk2", name="contact_3d")
