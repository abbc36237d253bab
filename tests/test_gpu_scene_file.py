"""Scenes read from TOML files, stepped on the GPU (needs a GPU).

- A scene file's fluid, solid and boundary particles step like the oracle's step of the same
  generated scene (normwise relative error <= 1e-5 per field after two steps), with the
  file's integrator.
- arp_scene_apply (scene upload + the file's integrator, time step and subticks in one C
  call) runs bitwise the same frames as the Python mirror built from the same file.
"""
import ctypes as C

import numpy as np
import pytest

from arp import config as K
from arp import tableau as T
from arp import wire

pytestmark = pytest.mark.gpu

FLUID_2D = """\
name = "Column 2D"
kernel_radius = 0.02
time_step = 0.0001
gravity = 9.81
artificial_viscocity = 0.05
integrator = "RK4"
subticks = 2
min = [-0.1, -0.1]
dim = [1.2, 0.8]

[boundaries.tank]
spacing = 0.01
region = {min = [-0.1, -0.1], dim = [1.2, 0.8], border = 0.03}

[objects.water]
density = 1000.0
speed_of_sound = 20.0
viscocity = 0.001
spacing = 0.01
region = {min = [0.0, 0.0], dim = [0.4, 0.3]}
velocity = [0.2, 0.0]

[objects.oil]
density = 800.0
speed_of_sound = 20.0
spacing = 0.01
region = {center = [0.7, 0.15], radii = [0.1, 0.1]}
water = {potential = "Tait", contact_strength = 0.5}
"""

SOLID_3D = """\
name = "Ball in a box"
kernel_radius = 0.04
time_step = 0.00002
gravity = 1.0
artificial_viscocity = 0.1
integrator = "Verlet"
min = [-0.3, -0.3, -0.3]
dim = [0.6, 0.6, 0.6]

[boundaries.box]
packing = 0.5
region = {min = [-0.3, -0.3, -0.3], dim = [0.6, 0.6, 0.6], border = 0.05}

[objects.ball]
density = 1.0
normal_stiffness = -40000.0
shear_stiffness = -40000.0
shear_dampening = -2.0
packing = 0.5
velocity = [0.3, 0.0, 0.0]
angular_velocity = [0.0, 0.0, 4.0]
region = {center = [0.0, 0.0, 0.0], radii = [0.1, 0.1, 0.1]}
box = {friction = 0.5}

[objects.water]
density = 1.0
speed_of_sound = 5.0
packing = 0.5
region = {min = [-0.2, -0.2, 0.12], dim = [0.3, 0.15, 0.1]}
"""


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _rel(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref))) / max(float(np.max(np.abs(ref))), 1e-300)


@pytest.mark.parametrize("name,text", [("fluid_2d", FLUID_2D), ("solid_3d", SOLID_3D)])
def test_scene_file_steps_like_the_oracle(torch_cuda, oracle_mod, name, text):
    from arp.sim import FluidSim

    sc = K.loads(text)
    print("case scene file", name, sc.integrator, sc.n, sc.nb, sc.ns, flush=True)
    assert sc.n > 50 and sc.nb > 50
    sim = FluidSim(sc)
    dt = sc.dt / sc.subticks
    sim.update(1)  # one frame = subticks steps of time_step / subticks
    got = sim.particles()
    o = oracle_mod.Oracle(sc)
    ref = o.step(o.state("f"), T.tableau(sc.integrator), dt, sc.subticks)
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], ref[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-5, (k, e)
    assert sim.time == pytest.approx(sc.dt, rel=1e-6)
    sim.close()


def test_scene_apply_matches_the_mirror(torch_cuda):
    from arp._lib import lib
    from arp.sim import FluidSim

    L = lib()
    text = FLUID_2D.encode()
    f = C.c_void_p()
    err = C.create_string_buffer(256)
    assert L.arp_scene_parse(text, len(text), None, C.byref(f), err, len(err)) == 0, err.value
    assert L.arp_scene_generate(f, 0) == 0
    ctx = C.c_void_p()
    assert L.arp_create(0, 0, C.byref(ctx)) == 0
    assert L.arp_scene_apply(ctx, f) == 0, L.arp_last_error(ctx)
    assert L.arp_update(ctx, 3) == 0, L.arp_last_error(ctx)
    sc = K.loads(FLUID_2D)
    a = wire.particles(sc.n)
    assert L.arp_download(ctx, a.ctypes.data_as(C.c_void_p), sc.n, None, 0) == 0
    L.arp_destroy(ctx)
    L.arp_scene_free(f)
    sim = FluidSim(sc)
    sim.update(3)
    b = sim.particles()
    for k in ("den", "pos", "vel"):
        assert np.array_equal(a[k], b[k]), k
    sim.close()
