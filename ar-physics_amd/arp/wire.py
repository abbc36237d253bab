"""std430 wire formats of the reference's GLSL structs, as numpy structured dtypes.

Byte-identical to ``include/arp.h`` and to the reference's GLSL/Rust structs:
``Particle``/``SolidParticle`` (src/soft_body/particle_state/mod.rs:24-35),
``Material`` (src/soft_body/material_region/material.rs:11-28),
``MatInteraction`` (src/soft_body/material_region/interaction.rs:10-15),
``AABB`` (src/soft_body/material_region/region.rs:11-14).
"""
import numpy as np

INVALID = 0xFFFFFFFF

PARTICLE = np.dtype(
    {
        "names": ["den", "mat", "solid_id", "pos", "vel"],
        "formats": ["<f4", "<u4", "<u4", ("<f4", 4), ("<f4", 4)],
        "offsets": [0, 4, 8, 16, 32],
        "itemsize": 48,
    }
)

SOLID = np.dtype(
    {
        "names": ["part_id", "ref_pos", "stress"],
        "formats": ["<u4", ("<f4", 4), ("<f4", 16)],
        "offsets": [0, 16, 32],
        "itemsize": 96,
    }
)

MATERIAL_FIELDS = [
    ("immobile", "<u4"),
    ("mass", "<f4"),
    ("start_den", "<f4"),
    ("state_eq", "<u4"),
    ("sound_speed", "<f4"),
    ("target_den", "<f4"),
    ("visc", "<f4"),
    ("bulk_visc", "<f4"),
    ("strain_order", "<i4"),
    ("normal_stiffness", "<f4"),
    ("shear_stiffness", "<f4"),
    ("normal_damp", "<f4"),
    ("shear_damp", "<f4"),
    ("plastic", "<u4"),
    ("yield_strength", "<f4"),
    ("work_hardening", "<f4"),
    ("work_hardening_exp", "<f4"),
    ("kinematic_hardening", "<f4"),
    ("thermal_softening", "<f4"),
    ("relaxation_time", "<f4"),
]
MATERIAL = np.dtype(MATERIAL_FIELDS)
assert MATERIAL.itemsize == 80

INTERACTION = np.dtype(
    [("potential", "<u4"), ("strength", "<f4"), ("radius", "<f4"), ("dampening", "<f4"), ("friction", "<f4")]
)
assert INTERACTION.itemsize == 20

AABB = np.dtype([("min", "<f4", 4), ("dim", "<f4", 4)])
assert AABB.itemsize == 32

MAX_STAGES = 6
TABLEAU = np.dtype(
    [
        ("stages", "<i4"),
        ("a", "<f4", (MAX_STAGES, MAX_STAGES)),
        ("b", "<f4", MAX_STAGES),
        ("c", "<f4", MAX_STAGES),
    ]
)

GRID_INFO = np.dtype(
    [
        ("sub", "<u4", 4),
        ("ncells", "<u8"),
        ("n_valid", "<u4", 3),
        ("n_overflow", "<u4", 3),
        ("n_nonfinite", "<u4", 3),
        ("occupied_rows", "<u4"),
    ],
    align=True,
)

# scene-file records (ArpSceneSettings / ArpSceneObject of include/arp.h)
SCENE_SETTINGS = np.dtype(
    [
        ("title", "S128"),
        ("integrator", "S32"),
        ("kernel_radius", "<f4"),
        ("time_step", "<f4"),
        ("artificial_viscosity", "<f4"),
        ("gravity", "<f4"),
        ("render_radius", "<f4"),
        ("subticks", "<u4"),
        ("lighting", "<u4"),
        ("bounds", AABB),
        ("n_objects", "<u4"),
        ("n_on_click", "<u4"),
        ("n_particles", "<u4"),
        ("n_boundary", "<u4"),
        ("n_solids", "<u4"),
    ]
)
assert SCENE_SETTINGS.itemsize == 240

SCENE_OBJECT = np.dtype(
    [
        ("name", "S64"),
        ("colors", "<f4", (3, 4)),
        ("packing", "<f4"),
        ("immobile", "<u4"),
        ("on_click", "<u4"),
        ("bound", AABB),
        ("material", MATERIAL),
        ("first", "<u4"),
        ("count", "<u4"),
        ("lattice_points", "<u4"),
    ]
)
assert SCENE_OBJECT.itemsize == 248

# state equations (material.rs:53-59) and contact potentials (interaction.rs:18-25)
EQ_ZERO, EQ_CONSTANT, EQ_IDEAL_GAS, EQ_TAIT = 0, 1, 2, 3
POT_ZERO, POT_CONSTANT, POT_LINEAR, POT_TAIT, POT_LENNARD_JONES, POT_HERTZ = range(6)


def particles(n):
    a = np.zeros(n, PARTICLE)
    a["solid_id"] = INVALID
    return a


def solids(n):
    return np.zeros(n, SOLID)


def material(**kw):
    m = np.zeros((), MATERIAL)
    for k, v in kw.items():
        m[k] = v
    return m


def interaction(potential=POT_ZERO, strength=0.0, radius=0.0, dampening=0.0, friction=0.0):
    i = np.zeros((), INTERACTION)
    i["potential"] = potential
    i["strength"] = strength
    i["radius"] = radius
    i["dampening"] = dampening
    i["friction"] = friction
    return i


def aabb(mn, dim):
    b = np.zeros((), AABB)
    b["min"][: len(mn)] = mn
    b["dim"][: len(dim)] = dim
    return b
