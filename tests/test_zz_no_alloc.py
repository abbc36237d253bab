"""arp_lincomb_noalloc (chunk partials in the library's fixed per-GPU scratch array) against
arp_lincomb (partials allocated per call), through the C ABI (needs a GPU).

Both run the same kernels in the same order, so the outputs must be equal bit for bit: two
calls in a row on one GPU, the second on a smaller scene, and a large scene at and beyond
the scratch limit."""
import numpy as np
import pytest

from arp import scene as S
from arp import wire
from helpers import box_scene, uniform_block

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solid_fluid_scene():
    sp = 0.02
    h = 1.3 * sp
    X = uniform_block(6, sp, 3, origin=(0.3, 0.3, 0.3))
    Fl = uniform_block(8, sp, 3, jitter=0.1, seed=3, origin=(0.25, 0.2, 0.25))
    Fl = Fl[~np.all(np.abs(Fl - X.mean(0)) < 4 * sp, axis=1)]
    P = np.concatenate([X, Fl])
    mask = np.concatenate([np.ones(len(X), bool), np.zeros(len(Fl), bool)])
    sm = S.solid_material(1000.0, -2e5, -1e5)
    sm["mass"] = 1000.0 * sp**3
    return box_scene(P, bounds=((0.0, 0.0, 0.0), (0.8, 0.8, 0.8)), h=h, solids_mask=mask, solid_mat=sm)


def _terms(torch, rng, sc, nterms, solids):
    from arp.sim import to_device

    td, sd = [], []
    for _ in range(nterms):
        p = sc.particles.copy()
        p["den"] = rng.normal(size=sc.n)
        p["pos"][:, :3] = rng.normal(size=(sc.n, 3))
        p["vel"][:, :3] = rng.normal(size=(sc.n, 3))
        td.append(to_device(p))
        if solids:
            s = sc.solids.copy()
            s["ref_pos"][:, :3] = rng.normal(size=(sc.ns, 3))
            s["stress"][:] = rng.normal(size=s["stress"].shape)
            sd.append(to_device(s))
    return td, (sd if solids else None)


def _both(torch, sim, sc, co, td, sd):
    from arp.sim import to_device

    outs = []
    for noalloc in (False, True):
        out = to_device(wire.particles(sc.n))
        sout = to_device(wire.solids(sc.ns)) if sd is not None else None
        sim.lincomb(co, td, out, solid_terms=sd, solid_out=sout, noalloc=noalloc)
        torch.cuda.synchronize()
        outs.append((out, sout))
    return outs


def test_noalloc_lincomb_equals_lincomb_twice(torch_cuda):
    from arp.sim import FluidSim

    torch = torch_cuda
    rng = np.random.default_rng(5)
    # first call: fluid + solids, 9 terms (two chunks of partials)
    sa = _solid_fluid_scene()
    assert sa.ns > 1
    print("case noalloc: solids scene n=%d ns=%d, 9 terms" % (sa.n, sa.ns), flush=True)
    sim_a = FluidSim(sa)
    td, sd = _terms(torch, rng, sa, 9, True)
    co = rng.normal(size=9).astype(np.float32)
    (o1, s1), (o2, s2) = _both(torch, sim_a, sa, co, td, sd)
    assert torch.equal(o1, o2) and torch.equal(s1, s2)
    assert o1.any().item() and s1.any().item()
    # second call, right after, on a smaller scene: 7 terms, no solids; the scratch array
    # still holds the first call's partials, which must not leak into this result
    sb = S.small_fluid_3d(n_side=5, box=(0.06, 0.06, 0.06))
    assert sb.n < sa.n
    print("case noalloc: fluid scene n=%d, 7 terms" % sb.n, flush=True)
    sim_b = FluidSim(sb)
    td, _ = _terms(torch, rng, sb, 7, False)
    co = rng.normal(size=7).astype(np.float32)
    (o1, _), (o2, _) = _both(torch, sim_b, sb, co, td, None)
    assert torch.equal(o1, o2) and o1.any().item()
    # few terms: one launch, no partials
    (o1, _), (o2, _) = _both(torch, sim_b, sb, co[:4], td[:4], None)
    assert torch.equal(o1, o2)
    sim_a.close()
    sim_b.close()


def test_noalloc_lincomb_scratch_limit(torch_cuda):
    from arp._lib import ArpError
    from arp.sim import FluidSim, to_device

    torch = torch_cuda
    sc = S.fluid_block_3d(600_000)
    print("case noalloc: limit, n=%d" % sc.n, flush=True)
    sim = FluidSim(sc)
    rng = np.random.default_rng(9)
    p = sc.particles.copy()
    p["vel"][:, :3] = rng.normal(size=(sc.n, 3))
    t = [to_device(p), to_device(sc.particles)]
    # 6 terms: two partials of 28.8 MB fit the 128 MiB array
    td = [t[k % 2] for k in range(6)]
    co = rng.normal(size=6).astype(np.float32)
    (o1, _), (o2, _) = _both(torch, sim, sc, co, td, None)
    assert torch.equal(o1, o2)
    # 21 terms: five partials (144 MB) do not fit; refused before anything runs
    td = [t[k % 2] for k in range(21)]
    out = to_device(wire.particles(sc.n))
    before = out.clone()
    with pytest.raises(ArpError, match="scratch"):
        sim.lincomb(np.ones(21, np.float32), td, out, noalloc=True)
    torch.cuda.synchronize()
    assert torch.equal(out, before)
    sim.close()
