"""Small scene factories shared by the tests."""
import math

import numpy as np

from arp import scene as S
from arp import wire


def box_scene(pos, vel=None, den=1000.0, bpos=None, bounds=((0, 0, 0), (1, 1, 1)), h=0.1, g=9.81,
              fluid=None, boundary_mat=None, solids_mask=None, solid_mat=None, bden=1.0,
              interactions=None, mats_extra=None):
    """A scene from explicit arrays: material 0 = boundary, 1 = fluid, 2 = solid (if any)."""
    pos = np.asarray(pos, np.float32)
    n = len(pos)
    p = wire.particles(n)
    p["pos"][:, : pos.shape[1]] = pos
    if vel is not None:
        vel = np.asarray(vel, np.float32)
        p["vel"][:, : vel.shape[1]] = vel
    p["den"] = den
    p["mat"] = 1
    bm = boundary_mat if boundary_mat is not None else S.immobile_material(0.0)
    if bm["mass"] == 0:
        bm["mass"] = 1e-3
    fm = fluid if fluid is not None else S.fluid_material(h, 0.02, density=1000.0, speed_of_sound=20.0)
    if fm["mass"] == 0:
        fm["mass"] = 1000.0 * (0.5 * h) ** 3
    mats = [bm, fm]
    if solids_mask is not None:
        sm = solid_mat if solid_mat is not None else S.solid_material(1000.0, -1e5, -5e4)
        if sm["mass"] == 0:
            sm["mass"] = fm["mass"]
        mats.append(sm)
        p["mat"][np.asarray(solids_mask)] = 2
    for m in mats_extra or ():  # further materials after these (ids 2, 3, ... or 3, ...)
        mats.append(m)
    mats = np.array(mats, wire.MATERIAL)
    if bpos is None:
        b = wire.particles(0)
    else:
        bpos = np.asarray(bpos, np.float32)
        b = wire.particles(len(bpos))
        b["pos"][:, : bpos.shape[1]] = bpos
        b["den"] = bden
        b["mat"] = 0
    inter = S.resolve_interactions(mats, h, interactions)
    lo, dim = bounds
    bd = wire.aabb(lo, dim)
    return S.Scene(p, b, mats, inter, bd, float(np.float32(h)), g, 0.02, dt=1e-4, integrator="rk4")


def brute_candidates(sc, pos=None):
    """O(N^2) candidate sets of the 27-cell stencil with the 128-slot cap, independent of the
    bucket traversal: cell coords from the reference formula, neighbours by |dc|_inf <= 1."""
    sub = np.zeros(3, np.int64)
    for i in range(3):
        q = np.float32(sc.bounds["dim"][i]) / np.float32(sc.h)
        sub[i] = int(max(np.float32(1.0), q)) if np.isfinite(q) else 1
    mn = sc.bounds["min"][:3].astype(np.float32)
    dm = sc.bounds["dim"][:3].astype(np.float32)

    def cells(P):
        P = np.asarray(P, np.float32)[:, :3]
        fin = np.all(np.isfinite(P), axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            t = (P - mn) / dm
        t = np.where(np.isnan(t), np.float32(0), t)
        t = np.clip(t, np.float32(0), np.float32(1))
        c = np.floor(t * sub.astype(np.float32)).astype(np.int64)
        c = np.minimum(c, sub - 1)
        return c, fin

    P = sc.particles["pos"] if pos is None else pos
    cp, fp = cells(P)
    cb, fb = cells(sc.boundary["pos"]) if sc.nb else (np.zeros((0, 3), np.int64), np.zeros(0, bool))
    has_ref = sc.particles["solid_id"] != wire.INVALID
    refpos = np.zeros((sc.n, 4), np.float32)
    refpos[has_ref] = sc.solids["ref_pos"][sc.particles["solid_id"][has_ref]]
    cr, fr = cells(refpos)
    fr &= has_ref

    def lin(c):
        return c[:, 0] + sub[0] * (c[:, 1] + sub[1] * c[:, 2])

    kb, kr, kp = lin(cb), lin(cr), lin(cp)
    room = {}
    keep_b = np.zeros(len(kb), bool)
    for key in np.unique(kb[fb]):
        ids = np.nonzero(fb & (kb == key))[0]
        keep_b[ids[:128]] = True
        room[key] = 128 - min(len(ids), 128)
    keep_r = np.zeros(sc.n, bool)
    for key in np.unique(kr[fr]):
        ids = np.nonzero(fr & (kr == key))[0]
        r = room.get(key, 128)
        keep_r[ids[:r]] = True
        room[key] = r - min(len(ids), r)
    keep_p = np.zeros(sc.n, bool)
    for key in np.unique(kp[fp]):
        ids = np.nonzero(fp & (kp == key))[0]
        keep_p[ids[: room.get(key, 128)]] = True
    cand = []
    for i in range(sc.n):
        if not keep_p[i]:
            cand.append(None)
            continue
        dp = np.abs(cp - cp[i]).max(1) <= 1
        db = np.abs(cb - cp[i]).max(1) <= 1 if sc.nb else np.zeros(0, bool)
        ids = list(np.nonzero(dp & keep_p)[0]) + [int(b) | 0x80000000 for b in np.nonzero(db & keep_b)[0]]
        cand.append(sorted(int(x) for x in ids))
    cell = np.where(keep_p, kp, 0xFFFFFFFF).astype(np.uint32)
    return cell, cand


def split_cand(cc, cand):
    out, off = [], 0
    for c in cc:
        out.append(list(cand[off: off + int(c)]))
        off += int(c)
    return out


def random_grid_scene(seed=0, dim=2):
    """Random particles with the grid's edge cases: NaN/inf, out of bounds (clamped),
    exactly at bounds.max, and a cluster that overflows the 128-slot bucket."""
    rng = np.random.default_rng(seed)
    h = 0.1
    n = 400
    pos = rng.uniform(-0.05, 1.05, (n, 3)).astype(np.float32)
    if dim == 2:
        pos[:, 2] = 0
    pos[:150] = np.float32(0.55) + rng.uniform(0, 0.02, (150, 3)).astype(np.float32)
    if dim == 2:
        pos[:150, 2] = 0
    pos[150] = [np.nan, 0.5, 0]
    pos[151] = [0.5, np.inf, 0]
    pos[152] = [1.0, 1.0, 0 if dim == 2 else 1.0]  # exactly bounds.max
    bpos = rng.uniform(0, 1, (60, 3)).astype(np.float32)
    bpos[:10] = np.float32(0.555)
    if dim == 2:
        bpos[:, 2] = 0
    dims = (1, 1, 0) if dim == 2 else (1, 1, 1)
    return box_scene(pos[:, :3], bpos=bpos, bounds=((0, 0, 0), dims), h=h)


def uniform_block(nside, spacing, dim=3, jitter=0.0, seed=0, origin=(0, 0, 0)):
    ax = [np.arange(nside)] * dim
    G = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, dim).astype(np.float64)
    P = (G + 0.5) * spacing + np.asarray(origin[:dim])
    if jitter:
        P += jitter * spacing * np.random.default_rng(seed).uniform(-1, 1, P.shape)
    out = np.zeros((len(P), 3))
    out[:, :dim] = P
    return out.astype(np.float32)


def kernel_closed_form(dim, h):
    return 4.0 / (math.pi * h**8) if dim == 2 else 315.0 / (64.0 * math.pi * h**9)


def deformed_solid_scene(dim=3, plastic=False, with_fluid=True, seed=0, order=2):
    """An elastic (optionally plastic) block of strain order `order` beside a jittered fluid
    on a boundary floor, the block deformed by a random affine map plus noise, every particle
    moving."""
    sp = 0.01
    h = 1.3 * sp
    nside = 6
    X = uniform_block(nside, sp, dim, origin=(0.3, 0.3, 0.3))
    parts = [X]
    mask = [np.ones(len(X), bool)]
    if with_fluid:
        Fl = uniform_block(8, sp, dim, jitter=0.1, seed=seed, origin=(0.25, 0.2, 0.25))
        Fl = Fl[~np.all(np.abs(Fl - X.mean(0)) < (nside / 2 + 1) * sp, axis=1)]
        parts.append(Fl)
        mask.append(np.zeros(len(Fl), bool))
    P = np.concatenate(parts)
    mask = np.concatenate(mask)
    sm = S.solid_material(1000.0, -2e5, -1e5, normal_damp=-20.0, shear_damp=-30.0, strain_order=order,
                          plastic=dict(yield_strength=50.0, relaxation_time=0.5) if plastic else None)
    sm["mass"] = 1000.0 * sp**dim
    fm = S.fluid_material(h, 0.02, density=1000.0, speed_of_sound=20.0, viscosity=0.01)
    fm["mass"] = 1000.0 * sp**dim
    B = uniform_block(12, sp, dim, origin=(0.2, 0.17, 0.2))
    B = B[B[:, 1] < 0.2]
    dims = (0.8, 0.8, 0.8) if dim == 3 else (0.8, 0.8, 0)
    sc = box_scene(P, bpos=B, bounds=((0.0, 0.0, 0.0), dims), h=h, fluid=fm, solids_mask=mask, solid_mat=sm)
    rng = np.random.default_rng(seed)
    # deform the solid: affine stretch + noise, and give everything a velocity field
    sid = sc.particles["solid_id"] != wire.INVALID
    c = sc.particles["pos"][sid, :dim].mean(0)
    A = np.eye(dim) + rng.normal(0, 0.03, (dim, dim))
    sc.particles["pos"][sid, :dim] = (sc.particles["pos"][sid, :dim] - c) @ A.T + c
    sc.particles["pos"][:, :dim] += rng.normal(0, 0.0005, (sc.n, dim)).astype(np.float32)
    sc.particles["vel"][:, :dim] = rng.normal(0, 0.05, (sc.n, dim)).astype(np.float32)
    sc.particles["den"] *= (1 + rng.uniform(-2e-3, 2e-3, sc.n)).astype(np.float32)
    if plastic:
        sc.solids["stress"][:, 0] = 1e-3
        sc.solids["stress"][:, 5] = -1e-3
    return sc


def self_contact_scene(seed=0, gap=0.3, outside=False):
    """Two 5^3 blocks of one elastic material, 0.4 m apart in the reference configuration
    (12.5 h, beyond the 8h self-contact cut) and side by side now with a gap of `gap`
    spacings along x, so self contact (solid.rs:426-450) acts between them; deformed and
    moving.  With `outside`, the grid ends inside the second block, whose outer particles
    then sit in clamped cells."""
    sp = 0.02
    h = 1.6 * sp
    A = uniform_block(5, sp, 3, jitter=0.1, seed=seed, origin=(0.1, 0.3, 0.3)).astype(np.float64)
    B = uniform_block(5, sp, 3, jitter=0.1, seed=seed + 1, origin=(0.1, 0.3, 0.3)).astype(np.float64)
    X = np.concatenate([A, B + np.array([0.0, 0.0, 0.4])])
    x = np.concatenate([A, B + np.array([(5 + gap) * sp, 0.0, 0.0])])
    rng = np.random.default_rng(seed)
    x = x + (x - x.mean(0)) @ (0.02 * rng.standard_normal((3, 3))).T + 0.03 * sp * rng.standard_normal(x.shape)
    sm = S.solid_material(1000.0, -4e5, -1e5, normal_damp=-30.0, shear_damp=-20.0)
    sm["mass"] = 1000.0 * sp**3
    bounds = ((0.0, 0.0, 0.0), (0.25 if outside else 1.0, 1.0, 1.0))
    sc = box_scene(X, bounds=bounds, h=h, solids_mask=np.ones(len(X), bool), solid_mat=sm)
    sc.particles["pos"][:, :3] = x.astype(np.float32)
    sc.particles["vel"][:, :3] = (0.1 * rng.standard_normal(x.shape)).astype(np.float32)
    sc.particles["den"] *= (1 + rng.uniform(-2e-3, 2e-3, sc.n)).astype(np.float32)
    return sc
