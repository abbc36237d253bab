"""Scene files: the TOML scenes of the reference's example runner, read and generated by the
native scene-file reader of libarp.so (include/arp.h, "scene files").

    sc = config.load("examples/dam_break.toml")     # Config::parse + FluidSim::new's lattice
    sim = FluidSim(sc)                              # its integrator, time_step and subticks

Reference: examples/test_mods/config.rs:298-474 (keys, defaults, material order),
src/soft_body/material_region/mod.rs:56-155 (lattice), src/soft_body/mod.rs:58-111
(materials, boundary split, interaction matrix).  Model paths ("model" keys) resolve against
``base_dir``, by default the working directory, as the reference's File::open does.
"""
import ctypes as C
import json
import os

import numpy as np

from . import wire
from ._lib import ArpError, lib
from .scene import Scene

__all__ = ["SceneFile", "load", "loads", "toml_to_json"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class SceneFile:
    """A parsed scene: settings and objects now, particles after ``generate``."""

    def __init__(self, text, base_dir=None):
        L = lib()
        if isinstance(text, str):
            text = text.encode()
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = L.arp_scene_parse(text, len(text), base_dir.encode() if base_dir else None, C.byref(self._h), err,
                               len(err))
        if rc != 0:
            self._h = None
            raise ArpError("scene file: " + err.value.decode(errors="replace"))

    def _check(self, rc):
        if rc != 0:
            raise ArpError("scene file error %d: %s" % (rc, lib().arp_scene_error(self._h).decode(errors="replace")))

    def settings(self):
        s = np.zeros((), wire.SCENE_SETTINGS)
        self._check(lib().arp_scene_settings(self._h, _p(s)))
        return s

    def object(self, i):
        o = np.zeros((), wire.SCENE_OBJECT)
        self._check(lib().arp_scene_object(self._h, i, _p(o)))
        return o

    def objects(self):
        st = self.settings()
        return [self.object(i) for i in range(int(st["n_objects"]) + int(st["n_on_click"]))]

    def generate(self, threads=0):
        """FluidSim::new's particle generation; returns a Scene."""
        L = lib()
        self._check(L.arp_scene_generate(self._h, threads))
        st = self.settings()
        n, nb, ns, nm = int(st["n_particles"]), int(st["n_boundary"]), int(st["n_solids"]), int(st["n_objects"])
        p = wire.particles(n)
        b = wire.particles(nb)
        so = wire.solids(ns)
        m = np.zeros(nm, wire.MATERIAL)
        it = np.zeros(nm * nm, wire.INTERACTION)
        self._check(L.arp_scene_arrays(self._h, _p(p), n, _p(so), ns, _p(b), nb, _p(m), nm, _p(it)))
        bounds = np.zeros((), wire.AABB)
        bounds["min"] = st["bounds"]["min"]
        bounds["dim"] = st["bounds"]["dim"]
        return Scene(p, b, m, it, bounds, float(st["kernel_radius"]), float(st["gravity"]),
                     float(st["artificial_viscosity"]), solids=so, dt=float(st["time_step"]),
                     subticks=int(st["subticks"]), integrator=bytes(st["integrator"][()]).decode().lower(),
                     name=bytes(st["title"][()]).decode(), objects=self.objects())

    def contains(self, obj, points):
        """Region::contains of object ``obj`` for an (n, 4) float32 array."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
        out = np.zeros(len(pts), np.uint8)
        self._check(lib().arp_scene_contains(self._h, obj, _p(pts), len(pts), _p(out)))
        return out.astype(bool)

    def close(self):
        if self._h:
            lib().arp_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def loads(text, base_dir=None, threads=0):
    """Scene of a TOML text."""
    f = SceneFile(text, base_dir)
    try:
        return f.generate(threads)
    finally:
        f.close()


def load(path, base_dir=None, threads=0):
    """Scene of a TOML file (model paths relative to base_dir, default the working
    directory)."""
    with open(path, "rb") as fh:
        text = fh.read()
    return loads(text, base_dir if base_dir is not None else os.getcwd(), threads)


def toml_to_json(text):
    """The TOML reader's view of a document, as Python data (floats stay floats)."""
    if isinstance(text, str):
        text = text.encode()
    L = lib()
    need = C.c_size_t(0)
    rc = L.arp_toml_to_json(text, len(text), None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    rc = L.arp_toml_to_json(text, len(text), buf, len(buf), C.byref(need))
    if rc != 0:
        raise ArpError(buf.value.decode(errors="replace"))
    return json.loads(buf.value.decode())
