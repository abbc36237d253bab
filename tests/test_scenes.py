"""The benchmark scenes themselves (CPU): the fluid-solid contact scene's ball gives every
solid particle an invertible corrected-gradient matrix, and the cantilever's clamp holds the
beam's root (oracle)."""
import numpy as np

from arp import scene as S
from arp import tableau as T
from arp import wire


def _corrected_gradient_matrices(X, h, V):
    """L_i = sum_j V grad W(X_j - X_i) (X_j - X_i)^T over |X_j - X_i| < h (strain.rs:101-118),
    with the poly6 gradient up to its positive normalisation."""
    from scipy.spatial import cKDTree

    tree = cKDTree(X)
    pairs = tree.query_pairs(h, output_type="ndarray")
    L = np.zeros((len(X), 3, 3))
    for a, b in ((pairs[:, 0], pairs[:, 1]), (pairs[:, 1], pairs[:, 0])):
        d = X[b] - X[a]
        q = h * h - (d * d).sum(1)
        g = -6.0 * (q * q)[:, None] * d
        np.add.at(L, a, V * g[:, :, None] * d[:, None, :])
    return L


def test_contact_ball_is_fully_supported():
    sc = S.contact_3d()
    solid = sc.particles["solid_id"] != wire.INVALID
    assert solid.sum() == 65536 and sc.n == 64**3 + 65536
    X = sc.particles["pos"][solid, :3].astype(np.float64)
    assert len(np.unique(X, axis=0)) == len(X)
    m = sc.materials[int(sc.particles["mat"][solid][0])]
    L = _corrected_gradient_matrices(X, sc.h, float(m["mass"] / m["start_den"]))
    cond = np.linalg.cond(L)
    print("cond(L) max %.2f" % cond.max())
    assert np.all(np.isfinite(cond)) and cond.max() < 10.0
    # the ball starts within the contact radius of the fluid and moving down
    fl = ~solid
    gap = X[:, 1].min() - sc.particles["pos"][fl, 1].max()
    assert 0 < gap < sc.h  # pairs interact within h
    assert np.allclose(sc.particles["vel"][solid, :3], [0.0, -2.0, 0.0])


def test_cantilever_clamp_holds_the_root(oracle_mod):
    """A 4 x 4 x 24 beam for 8 ms: its free end falls (the reference doubles pos' and gravity
    of elastic particles, so free fall is -2 g t^2, and a beam this soft is still nearly in
    free fall), its clamped root far less; nothing moves across the plates."""
    sc = S.cantilever_3d(nx=4, ny=4, nz=24)
    o = oracle_mod.Oracle(sc)
    st = o.state("f")
    p0 = st["pos"][:, :3].copy()
    st = o.step(st, T.tableau("rk4"), 1e-5, 800)
    d = st["pos"][:, :3] - p0
    assert np.isfinite(d).all()
    z = p0[:, 2]
    root = z < 8 * 0.0025
    tip = z > z.max() - 0.003
    fall = 2 * 9.81 * 8e-3**2
    print("root dy %.3e tip dy %.3e free fall %.3e" % (d[root, 1].mean(), d[tip, 1].mean(), -fall))
    assert d[tip, 1].mean() < -0.8 * fall
    assert abs(d[root, 1].mean()) < 0.6 * abs(d[tip, 1].mean())
    assert np.abs(d[root][:, [0, 2]]).max() < 0.25 * 0.0025  # the grip squeezes the corners a little
