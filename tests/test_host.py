"""Host-side checks of the C ABI library (no GPU): exports, wire layouts, FluidSim::new
derivations."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from arp import _lib
from arp import scene as S
from arp import wire

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "arp.h")).read()
    return sorted(set(re.findall(r"\b(arp_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SYMBOLS)


def test_wire_sizes_match_header():
    assert wire.PARTICLE.itemsize == 48
    assert wire.SOLID.itemsize == 96
    assert wire.MATERIAL.itemsize == 80
    assert wire.INTERACTION.itemsize == 20
    assert wire.AABB.itemsize == 32
    assert wire.PARTICLE.fields["pos"][1] == 16 and wire.PARTICLE.fields["vel"][1] == 32
    assert wire.SOLID.fields["ref_pos"][1] == 16 and wire.SOLID.fields["stress"][1] == 32


@pytest.mark.parametrize("dim,h", [(2, 0.013), (3, 0.0065), (3, 1.0)])
def test_norm_const_matches_oracle(oracle_mod, dim, h):
    assert _lib.lib().arp_norm_const(dim, h) == oracle_mod.norm_const(dim, h, "f")


def test_default_and_resolved_interactions():
    L = _lib.lib()
    h = np.float32(0.013)
    wall = S.immobile_material(0.25)
    water = S.fluid_material(h, 0.02, density=1000.0, speed_of_sound=20.0, viscosity=0.1)
    solid = S.solid_material(1.0, -4e5, -5e4)
    mats = np.array([wall, water, solid], wire.MATERIAL)
    out = np.zeros(9, wire.INTERACTION)
    cfg = np.zeros(9, wire.INTERACTION)
    present = np.zeros(9, np.uint8)
    cfg[1 * 3 + 2] = wire.interaction(wire.POT_TAIT, 0.05, 0.026, 0.1, 0.2)
    present[1 * 3 + 2] = 1
    L.arp_resolve_interactions(3, mats.ctypes.data_as(C.c_void_p), cfg.ctypes.data_as(C.c_void_p),
                               present.ctypes.data_as(C.c_void_p), h, out.ctypes.data_as(C.c_void_p))
    ref = S.resolve_interactions(mats, h, {(1, 2): cfg[5]})
    assert out.tobytes() == ref.tobytes()
    # fluid-wall pair: linear, strength 1e4, radius 2h, friction visc sum (interaction.rs:42-50)
    assert out[1]["potential"] == wire.POT_LINEAR and out[1]["radius"] == np.float32(2 * h)
    assert out[1]["friction"] == np.float32(0.35)
    # same-material pairs: zero interaction; config applies symmetrically
    assert out[4]["potential"] == wire.POT_ZERO
    assert out[7].tobytes() == out[5].tobytes()


def test_tableau_rejects_implicit():
    L = _lib.lib()
    t = np.zeros((), wire.TABLEAU)
    assert L.arp_tableau_by_name(b"nonsense", t.ctypes.data_as(C.c_void_p)) != 0


def test_graft_entry_build_is_importable():
    import importlib.util

    spec = importlib.util.spec_from_file_location("ge", os.path.join(ROOT, "__graft_entry__.py"))
    ge = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ge)
    assert callable(ge.build) and callable(ge.smoke)


def test_neighbor_list_size_limit():
    """The list walks address entries with 32-bit byte offsets: lists are kept only while
    cap * (n + 16) * 4 bytes stay below 4 GiB (entry k of particle i at k * (n + 16) + i)."""
    L = _lib.lib()
    for cap in (8, 48, 1024):
        nmax = -(-(1 << 32) // (4 * cap)) - 16  # first n whose entries reach 2^32 bytes
        assert L.arp_neighbor_lists_fit(cap, nmax - 1) == 1
        assert L.arp_neighbor_lists_fit(cap, nmax) == 0
    assert L.arp_neighbor_lists_fit(0, 1 << 20) == 1  # default 48 entries at the 1 M column
    assert L.arp_neighbor_lists_fit(1024, 1 << 20) == 0  # 4 GiB of lists: not kept
