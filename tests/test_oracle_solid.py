"""Pins the oracle's solid_forces (solid.rs:276-524) and its strain input (strain.rs:71-181)
with what the mathematics fixes, independently of the oracle's own code:

- rigid motion (rotation + translation, rigid-body velocity field) is force free: F = R
  exactly through the corrected gradient, so Green-Lagrange strain, its rate, the elastic
  and hourglass forces vanish and a = -g;
- every pair term (elastic, hourglass, self contact) is antisymmetric, so the forces sum to
  zero over a solid-only scene;
- the elastic force is the gradient of the St Venant-Kirchhoff energy sum_i V_i psi(F_i)
  of the corrected-gradient F (differentiated here by torch autograd, not by the oracle's
  formula), with the reference's sign: its stiffnesses enter negated (f = +dE/dx);
- a uniform stretch / shear gives the Hooke traction across a cut plane (continuum limit);
- the plastic rate switches on exactly at |S| = A, with the yield stress A whatever the
  plastic strain (the integer (dim-1)/dim quirk, solid.rs:212-215), and equals
  (|S| - A) / T * S / |S| above it;
- self contact of a pair far apart in the reference configuration and near now is
  -5000 sqrt((8h - l) h / 8) r / l (solid.rs:426-450), over the 27-cell stencil.

All checks run the FP64 oracle (the FP32 build follows the same source)."""
import math

import numpy as np
import pytest

from arp import scene as S
from arp import wire
from helpers import box_scene, uniform_block

G = float(np.float32(9.81))  # the scene uniform is f32


def _solid_scene(P, h, mat, bounds=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))):
    P = np.asarray(P, np.float32)
    mat = mat.copy()
    if mat["mass"] == 0:
        mat["mass"] = 1.0
    return box_scene(P, bounds=bounds, h=h, g=G, solids_mask=np.ones(len(P), bool), solid_mat=mat)


def _state(o, X, x, v=None, den=None, eps=None):
    """FP64 state: reference X, current x, velocity v, density, plastic strain eps (mat4s,
    row-major [s, row, col])."""
    st = o.state("d")
    n = len(X)
    st["ref"][:, :3] = X
    st["pos"][:, :3] = x
    st["vel"][:] = 0
    if v is not None:
        st["vel"][:, :3] = v
    if den is not None:
        st["den"][:] = den
    if eps is not None:
        st["stress"][:] = np.transpose(eps, (0, 2, 1)).reshape(n, 16)  # column-major
    return st


def _forces(o, st):
    """rho_i (a_i + g y) of solid_forces alone (solid.rs:492-494: vel' = F / rho - g y)."""
    d = o.compute_forces(st, which="solid")
    a = d["vel"][:, :3].copy()
    a[:, 1] += G
    return d, st["den"][:, None] * a


def _rot(axis, th):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


SP = 0.02
H = 1.6 * SP


def _mat(lam=-4e5, mu=-1e5, nd=0.0, sd=0.0, plastic=None):
    m = S.solid_material(1000.0, lam, mu, normal_damp=nd, shear_damp=sd, plastic=plastic)
    m["mass"] = 1000.0 * SP**3  # density x lattice volume
    return m


def test_rigid_motion_is_force_free(oracle_mod):
    X = uniform_block(6, SP, 3, jitter=0.15, seed=4, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    sc = _solid_scene(X, H, _mat(nd=-30.0, sd=-20.0))
    o = oracle_mod.Oracle(sc)
    R = _rot((1.0, 2.0, -0.5), 0.6)
    c = X.mean(0)
    x = (X - c) @ R.T + c + np.array([0.01, -0.02, 0.015])
    w = np.array([0.3, -1.1, 0.7])  # rigid angular velocity: v = w x (x - c) + u
    v = np.cross(w, x - c) + np.array([0.1, 0.0, -0.2])
    st = _state(o, X, x, v)
    s = o.strain(st)
    F = np.transpose(s[:, 1], (0, 2, 1))[:, :3, :3]
    assert np.allclose(F, R, atol=1e-12)  # every particle: the correction makes F exact
    d, f = _forces(o, st)
    # force scale of a 1 % strain for comparison
    scale = 4e5 * 0.01 * SP**2
    assert np.abs(f).max() < 1e-8 * scale, np.abs(f).max()
    assert np.array_equal(d["pos"][:, :3], st["vel"][:, :3])
    assert np.all(d["den"] == 0) and np.all(d["stress"] == 0) and np.all(d["ref"] == 0)
    # a non-rigid velocity field makes the damping stress act (the check is not vacuous)
    st2 = _state(o, X, x, v + 0.05 * (x - c) @ np.diag([1.0, -0.5, 0.2]).T)
    _, f2 = _forces(o, st2)
    assert np.abs(f2).max() > 1e-3 * scale


def _two_blocks(seed, n_side=5, gap=2.0):
    """Two blocks of one material: far apart in the reference configuration (more than 8h),
    side by side now with a gap of `gap` spacings (so self contact acts)."""
    A = uniform_block(n_side, SP, 3, jitter=0.1, seed=seed, origin=(0.1, 0.3, 0.3)).astype(np.float64)
    B = uniform_block(n_side, SP, 3, jitter=0.1, seed=seed + 1, origin=(0.1, 0.3, 0.3)).astype(np.float64)
    X = np.concatenate([A, B + np.array([0.0, 0.0, 0.4])])  # 0.4 m = 12.5 h apart
    x = np.concatenate([A, B + np.array([(n_side + gap) * SP, 0.0, 0.0])])
    return X, x


def test_pair_terms_sum_to_zero(oracle_mod):
    """Elastic, hourglass and self-contact pair terms are antisymmetric (solid.rs:343-450):
    sum_i rho_i (a_i + g y) = 0 for a deformed, moving, solid-only scene."""
    rng = np.random.default_rng(8)
    X, x = _two_blocks(8, gap=0.3)
    n = len(X)
    x = x + (x - x.mean(0)) @ (0.03 * rng.standard_normal((3, 3))).T + 0.05 * SP * rng.standard_normal((n, 3))
    v = 0.2 * rng.standard_normal((n, 3))
    den = 1000.0 * (1 + 0.01 * rng.standard_normal(n))
    e = 1e-3 * rng.standard_normal((n, 3, 3))
    eps = np.zeros((n, 4, 4))
    eps[:, :3, :3] = e + np.transpose(e, (0, 2, 1))
    sc = _solid_scene(X, H, _mat(nd=-30.0, sd=-20.0))
    o = oracle_mod.Oracle(sc)
    st = _state(o, X, x, v, den, eps)
    _, f = _forces(o, st)
    tot = f.sum(0)
    mag = np.abs(f).sum(0)
    assert np.all(np.abs(tot) <= 1e-10 * mag), (tot, mag)
    # self contact is part of it: the only force between the blocks, it pushes them apart
    push = f[: n // 2].sum(0)
    assert push[0] < -1e-3 * mag[0], (push, mag)
    # the same blocks placed as in the reference configuration (no contact)
    st_far = _state(o, X, X, v, den, eps)
    _, f_far = _forces(o, st_far)
    assert np.abs(f_far.sum(0)).max() <= 1e-10 * np.abs(f_far).sum()
    assert np.abs(f_far[: n // 2].sum(0)).max() <= 1e-10 * np.abs(f_far).sum()
    # the pair weights V1 V2 = m^2 / (rho_i rho_j) are symmetric for any densities
    st2 = _state(o, X, x, v, den * np.where(np.arange(n) < n // 2, 1.3, 0.9), eps)
    _, f2 = _forces(o, st2)
    assert np.all(np.abs(f2.sum(0)) <= 1e-10 * np.abs(f2).sum(0))


def _svk_energy_grad(X, x, V, lam, mu, h, eps):
    """dE/dx of E = sum_i V psi(E_i + eps_i), psi(E) = lam/2 tr(E)^2 + mu E:E, with
    F_i = sum_j V (x_j - x_i) (L_i^-1 grad W(X_j - X_i))^T and L_i = sum_j V grad W(X_j - X_i)
    (X_j - X_i)^T over |X_j - X_i| < h (torch autograd, float64).  The poly6 normalisation
    cancels in F and in the force, so W = (h^2 - r^2)^3."""
    import torch

    Xt = torch.tensor(X, dtype=torch.float64)
    dX = Xt[None, :, :] - Xt[:, None, :]  # [i, j] = X_j - X_i
    r2 = (dX**2).sum(-1)
    q = torch.clamp(h * h - r2, min=0.0)
    gw = -6.0 * q[..., None] ** 2 * dX  # grad W(X_j - X_i)
    Lm = V * torch.einsum("ija,ijb->iab", gw, dX)
    Linv = torch.linalg.inv(Lm)
    gr = torch.einsum("iab,ijb->ija", Linv, gw)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    dx = xt[None, :, :] - xt[:, None, :]
    F = V * torch.einsum("ija,ijb->iab", dx, gr)
    E = 0.5 * (F.transpose(1, 2) @ F - torch.eye(3, dtype=torch.float64)) + torch.tensor(eps, dtype=torch.float64)
    tr = E.diagonal(dim1=1, dim2=2).sum(-1)
    psi = 0.5 * lam * tr**2 + mu * (E * E).sum((1, 2))
    (V * psi).sum().backward()
    return xt.grad.numpy()


def test_elastic_force_is_energy_gradient(oracle_mod):
    rng = np.random.default_rng(3)
    X = uniform_block(5, SP, 3, jitter=0.15, seed=6, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    n = len(X)
    c = X.mean(0)
    x = c + (X - c) @ (np.eye(3) + 0.04 * rng.standard_normal((3, 3))).T + 0.04 * SP * rng.standard_normal((n, 3))
    e = 2e-3 * rng.standard_normal((n, 3, 3))
    e = e + np.transpose(e, (0, 2, 1))
    eps = np.zeros((n, 4, 4))
    eps[:, :3, :3] = e
    lam, mu = -4e5, -1e5
    mat = _mat(lam, mu)
    mat["mass"] = 1000.0 * SP**3
    V = float(np.float32(mat["mass"])) / 1000.0  # m / start_den; den = start_den below
    o = oracle_mod.Oracle(_solid_scene(X, H, mat))
    tiny = _mat(lam * 1e-12, mu * 1e-12)
    tiny["mass"] = mat["mass"]
    o0 = oracle_mod.Oracle(_solid_scene(X, H, tiny))
    st = _state(o, X, x, den=np.full(n, 1000.0), eps=eps)
    _, f = _forces(o, st)
    _, f0 = _forces(o0, _state(o0, X, x, den=np.full(n, 1000.0), eps=eps))
    fe = f - f0  # the elastic force alone: hourglass and contact do not depend on lam, mu
    g = _svk_energy_grad(X, x, V, lam, mu, float(np.float32(H)), e)
    assert np.abs(g).max() > 0
    err = np.abs(fe - g).max() / np.abs(g).max()
    assert err < 1e-9, err
    # the physical (negated) stiffness would give the opposite sign
    assert np.abs(fe + g).max() > np.abs(g).max()


def test_hourglass_restores_a_displaced_particle(oracle_mod):
    """The hourglass term (solid.rs:430-443) penalises the part of a pair's current offset
    that the mean deformation gradient does not explain: with the stiffnesses near zero, an
    interior particle moved off a rigidly placed lattice is pulled back (f . d < 0, mostly
    along -d), and the forces balance; an affine
    placement (same F everywhere) gives no hourglass force at all."""
    X = uniform_block(7, SP, 3, jitter=0.1, seed=12, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    n = len(X)
    c = X.mean(0)
    k = int(np.argmin(np.linalg.norm(X - c, axis=1)))
    mat = _mat(-4e5 * 1e-12, -1e5 * 1e-12)
    o = oracle_mod.Oracle(_solid_scene(X, H, mat))
    dirn = np.array([0.6, -0.3, 0.74])
    dirn /= np.linalg.norm(dirn)
    for amp in (0.02, 0.04):
        x = X.copy()
        x[k] += amp * SP * dirn
        _, f = _forces(o, _state(o, X, x, den=np.full(n, 1000.0)))
        fk = f[k]
        assert fk @ dirn < 0 and abs(fk @ dirn) > 0.9 * np.linalg.norm(fk), fk
        assert np.abs(f.sum(0)).max() < 1e-10 * np.abs(f).sum()
        if amp == 0.02:
            f1 = fk @ dirn
    assert 1.5 < (fk @ dirn) / f1 < 2.5  # about linear in the displacement
    A = np.eye(3) + 0.05 * np.random.default_rng(1).standard_normal((3, 3))
    _, f = _forces(o, _state(o, X, c + (X - c) @ A.T, den=np.full(n, 1000.0)))
    assert np.abs(f).max() < 1e-9 * abs(f1)


@pytest.mark.parametrize("mode", ["stretch", "shear"])
def test_uniform_deformation_traction(oracle_mod, mode):
    """Affine x = A X: the hourglass vanishes and the summed force on the particles beyond a
    cut plane x > x0 equals the Hooke (St Venant-Kirchhoff) traction P n area of the cut, with
    the reference's sign (P of the negated stiffnesses); stretch pins lam + 2 mu, shear pins
    mu.  On this regular lattice the corrected-gradient pair sums reproduce the continuum
    traction to about 1e-8."""
    nx, ny, nz = 16, 12, 12
    ax = [np.arange(k) for k in (nx, ny, nz)]
    X = (np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3) + 0.5) * SP + 0.2
    E_, nu = 1e6, 0.4
    lam = -E_ * nu / ((1 + nu) * (1 - 2 * nu))
    mu = -E_ / (2 * (1 + nu))
    mat = _mat(lam, mu)
    mat["mass"] = 1000.0 * SP**3
    o = oracle_mod.Oracle(_solid_scene(X, H, mat, bounds=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))))
    gam = 1e-4
    A = np.eye(3)
    if mode == "stretch":
        A[0, 0] += gam
    else:
        A[1, 0] += gam  # x-faces slide along y
    c = X.mean(0)
    x = c + (X - c) @ A.T
    st = _state(o, X, x, den=np.full(len(X), 1000.0))
    _, f = _forces(o, st)
    right = X[:, 0] > c[0]
    got = f[right].sum(0)
    Eg = 0.5 * (A.T @ A - np.eye(3))
    Sg = lam * np.trace(Eg) * np.eye(3) + 2 * mu * Eg
    P = A @ Sg
    area = (ny * SP) * (nz * SP)
    expect = P @ np.array([1.0, 0.0, 0.0]) * area
    k = 0 if mode == "stretch" else 1
    print(mode, got, expect, flush=True)
    assert abs(got[k] - expect[k]) < 1e-6 * abs(expect[k]), (got, expect)
    # swapping lam and mu, dropping the factor 2 or the sign would miss by far more
    other = (mu * np.trace(Eg) * np.eye(3) + 2 * lam * Eg)
    assert abs((A @ other @ [1, 0, 0])[k] * area - got[k]) > 0.3 * abs(expect[k])


@pytest.mark.parametrize("side", [-1, 1])
def test_plastic_rate_threshold(oracle_mod, side):
    """solid.rs:499-518 with eq_plastic_strain_sq = 0 (integer (dim-1)/dim): the yield
    stress is A for any plastic strain, however large B; below A no flow, above it
    (|S| - A) / T * S / |S|.  S is the 4x4 PK2 stress of the reference (hooke uses the full
    identity, so its (3,3) entry is lam tr E)."""
    X = uniform_block(6, SP, 3, jitter=0.1, seed=2, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    n = len(X)
    c = X.mean(0)
    A = np.array([[1.002, 0.001, 0.0], [0.0005, 0.999, 0.0], [0.0, 0.0, 1.001]])
    x = c + (X - c) @ A.T
    ep = np.zeros((4, 4))
    ep[:3, :3] = [[2e-3, 1e-3, 0.0], [1e-3, -1e-3, 0.0], [0.0, 0.0, -1e-3]]
    lam, mu = -4e5, -1e5
    E4 = np.zeros((4, 4))
    E4[:3, :3] = 0.5 * (A.T @ A - np.eye(3))
    E4 += ep
    S4 = lam * np.trace(E4[:3, :3]) * np.eye(4) + 2 * mu * E4
    sn = math.sqrt((S4 * S4).sum())
    Ay = sn / (1 + side * 1e-3)
    T = 0.5
    mat = _mat(lam, mu, plastic=dict(yield_strength=Ay, work_hardening=1e9, work_hardening_exp=2.0,
                                     relaxation_time=T))
    mat["mass"] = 1000.0 * SP**3
    o = oracle_mod.Oracle(_solid_scene(X, H, mat))
    st = _state(o, X, x, eps=np.broadcast_to(ep, (n, 4, 4)))
    d = o.compute_forces(st, which="solid")
    rate = np.transpose(d["stress"].reshape(n, 4, 4), (0, 2, 1))  # row-major
    if side < 0:
        assert np.all(rate == 0)
    else:
        Ayf = float(np.float32(Ay))
        expect = (sn - Ayf) / T * S4 / sn
        assert np.allclose(rate, expect, rtol=1e-6, atol=1e-9 * np.abs(expect).max())


def test_self_contact_closed_form(oracle_mod):
    """Two rigid clusters of one material, 12.5 h apart in the reference configuration and
    0.3 spacings apart now: each pair within 8h whose cells are neighbours pushes with
    -5000 sqrt((8h - l) h / 8) r / l (solid.rs:446-450); nothing else acts."""
    X, x = _two_blocks(1, n_side=3, gap=0.3)
    n = len(X)
    mat = _mat()
    mat["mass"] = 1000.0 * SP**3
    sc = _solid_scene(X, H, mat)
    o = oracle_mod.Oracle(sc)
    den = np.full(n, 1000.0)
    st = _state(o, X, x, den=den)
    _, f = _forces(o, st)
    h = float(np.float32(H))
    sub = np.array([int(np.float32(1.0) / np.float32(h))] * 3)  # trunc(dim / h), bounds [0, 1]^3
    cell = np.floor(np.clip(x, 0, 1) * sub).astype(int)
    blk = np.arange(n) >= n // 2
    expect = np.zeros((n, 3))
    npairs = 0
    for i in range(n):
        for j in range(n):
            if blk[i] == blk[j] or np.abs(cell[i] - cell[j]).max() > 1:
                continue
            r = x[j] - x[i]
            l = np.linalg.norm(r)
            if l >= 8 * h:
                continue
            npairs += 1
            expect[i] += -5000.0 * math.sqrt((8 * h - l) * h / 8) * r / l
    assert npairs > 20
    assert np.allclose(f, expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max())


def _cm(M):
    return np.ascontiguousarray(np.asarray(M, np.float32).T.reshape(16))  # column-major


def test_seth_hill_orders(oracle_mod):
    """seth_hill / seth_hill_rate (solid.rs:147-199) for every order the scenes use:
    2m = 1 (Denman-Beavers square root, 15 iterations) is the Biot strain U - I, 2m = 3 is
    (C^(3/2) - I)/3, 2m = -2 is (I - C^-1)/2, 2m = 4 is (C^2 - I)/4 (scipy sqrtm / inv);
    the rate of order 2 is sym(F'^T F), of order 4 (F'^T F + F^T F') C-symmetrised / 4 as the
    reference sums dC C + C dC; of order 0 it is dC C^-1 / 2 (the reference's formula)."""
    import ctypes as C

    from scipy.linalg import sqrtm

    L = oracle_mod.lib()
    cp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rng = np.random.default_rng(21)
    F = np.eye(4)
    F[:3, :3] += rng.normal(0, 0.05, (3, 3))
    F[3, 3] = 0
    Fd = np.zeros((4, 4))
    Fd[:3, :3] = rng.normal(0, 0.3, (3, 3))
    C3 = F[:3, :3].T @ F[:3, :3]
    dC = Fd[:3, :3].T @ F[:3, :3]
    dC = dC + dC.T
    I3 = np.eye(3)
    out = np.zeros(16, np.float32)

    def sh(two_m):
        L.orc_seth_hill_f(cp(_cm(F)), two_m, 3, cp(out))
        return out.reshape(4, 4).T.astype(np.float64)

    def shr(two_m):
        L.orc_seth_hill_rate_f(cp(_cm(F)), cp(_cm(Fd)), two_m, 3, cp(out))
        return out.reshape(4, 4).T.astype(np.float64)

    U = sqrtm(C3).real
    expect = {1: U - I3, 3: (C3 @ U - I3) / 3, -2: (np.linalg.inv(C3) - I3) / -2, 4: (C3 @ C3 - I3) / 4,
              2: (C3 - I3) / 2}
    for two_m, e in expect.items():
        got = sh(two_m)
        assert np.allclose(got[:3, :3], e, atol=2e-6), (two_m, got[:3, :3], e)
        assert np.allclose(got[3, :], 0, atol=1e-7) and np.allclose(got[:, 3], 0, atol=1e-7)
    assert np.allclose(shr(2)[:3, :3], dC / 2, atol=2e-6)
    assert np.allclose(shr(4)[:3, :3], (dC @ C3 + C3 @ dC) / 4, atol=2e-6)
    assert np.allclose(shr(0)[:3, :3], 0.5 * dC @ np.linalg.inv(C3), atol=2e-6)
