"""Scene files (CPU): the native TOML reader and scene generator of libarp.so against the
numpy restatement in oracle/scene_ref.py, and that restatement against closed forms.

- Oracle pinning: region membership on and around known boundaries (boxes, ellipsoids,
  shells, CSG, scale / translation / axis rotations, a cube mesh and its surface), the
  lattice's spacings, offsets and counts, and the lattice mass.
- TOML reader: the parsed tree equals tomli's on a syntax corpus; both reject the same
  malformed documents.
- Scenes: particles, boundary, solids, materials and the interaction matrix are bit-exact
  with the oracle on a synthetic scene that uses every key family, and on the reference's
  own example scenes when the reference tree is present (models: on a sample of lattice
  points, the brute-force mesh tests being slow in numpy).
- The reference's error behaviour (a panic there, an error here) for wrong types, unknown
  names and missing regions or models.
"""
import math
import os
import struct

import numpy as np
import pytest

from arp import config as K
from arp._lib import ArpError
from oracle import scene_ref as R

tomli = pytest.importorskip("tomli")
f32 = np.float32
REF_EXAMPLES = "/root/reference/examples"


# ---------------------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------------------
def cube_tris(lo=0.0, hi=1.0):
    """The 12 outward triangles of the cube [lo, hi]^3 as (normal, v0, v1, v2)."""
    c = [[lo if (k >> b) & 1 == 0 else hi for b in range(3)] for k in range(8)]
    quads = [((0, 0, -1), (0, 2, 3, 1)), ((0, 0, 1), (4, 5, 7, 6)), ((0, -1, 0), (0, 1, 5, 4)),
             ((0, 1, 0), (2, 6, 7, 3)), ((-1, 0, 0), (0, 4, 6, 2)), ((1, 0, 0), (1, 3, 7, 5))]
    tris = []
    for n, (a, b, cc, d) in quads:
        tris.append((n, c[a], c[b], c[cc]))
        tris.append((n, c[a], c[cc], c[d]))
    return tris


def write_stl(path, tris, ascii_=False):
    if ascii_:
        with open(path, "w") as f:
            f.write("solid cube\n")
            for n, a, b, c in tris:
                f.write("  facet normal %r %r %r\n    outer loop\n" % tuple(map(float, n)))
                for v in (a, b, c):
                    f.write("      vertex %r %r %r\n" % tuple(map(float, v)))
                f.write("    endloop\n  endfacet\n")
            f.write("endsolid cube\n")
    else:
        with open(path, "wb") as f:
            f.write(b"binary cube".ljust(80, b" "))
            f.write(struct.pack("<I", len(tris)))
            for n, a, b, c in tris:
                f.write(struct.pack("<12fH", *n, *a, *b, *c, 0))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_same_scene(sc, ref):
    P, B = ref["particles"], ref["boundary"]
    for k in ("den", "mat", "solid_id", "pos", "vel"):
        assert sc.particles[k].shape == P[k].shape, ("particles", k, sc.particles[k].shape, P[k].shape)
        assert np.array_equal(bits(sc.particles[k]), bits(P[k])), ("particles", k)
    for k in ("den", "mat", "pos", "vel"):
        assert sc.boundary[k].shape == B[k].shape, ("boundary", k)
        assert np.array_equal(bits(sc.boundary[k]), bits(B[k])), ("boundary", k)
    assert np.array_equal(sc.solids["part_id"], ref["solids"]["part_id"])
    assert np.array_equal(bits(sc.solids["ref_pos"]), bits(ref["solids"]["ref_pos"]))
    assert len(sc.materials) == len(ref["materials"])
    for i, m in enumerate(ref["materials"]):
        for k, v in m.items():
            assert np.asarray(sc.materials[k][i]).tobytes() == np.asarray(v, sc.materials.dtype[k]).tobytes(), \
                ("material", i, k, sc.materials[k][i], v)
    nm = len(ref["materials"])
    for i in range(nm):
        for j in range(nm):
            for k, v in ref["inter"][i][j].items():
                got = sc.interactions[k][i * nm + j]
                assert np.asarray(got).tobytes() == np.asarray(v, sc.interactions.dtype[k]).tobytes(), \
                    ("interaction", i, j, k, got, v)


def oracle_scene(text, base_dir=None):
    return R.generate(R.parse(tomli.loads(text), base_dir))


# ---------------------------------------------------------------------------------------
# oracle pinned by closed forms
# ---------------------------------------------------------------------------------------
def test_lattice_mass_and_count():
    # staggered lattice: rows of s, s/sqrt(2); mass = density * area / points tested
    L = R.lattice(R.AABB([0, 0], [1.0, 1.0]), 0.1)
    # f32 accumulation of 0.1 passes 1.0 after 10 points (offsets 0 and 0.05 alike);
    # 15 rows 0.0707 apart fit in the unit height
    assert len(L) == 10 * 15
    assert R.lattice_mass(1000.0, R.AABB([0, 0], [1.0, 1.0]), len(L)) == f32(1000.0 / 150)
    xs = np.sort(np.unique(L[:, 0]))
    assert len(xs) == 20 and xs[1] == f32(0.05)
    ys = np.unique(L[:, 1])
    assert np.allclose(np.diff(ys), 0.1 / math.sqrt(2), rtol=1e-5)


def test_lattice_3d_layers():
    h = 0.1
    L = R.lattice(R.AABB([0, 0, 0], [0.5, 0.5, 0.5]), h)
    zs = np.unique(L[:, 2])
    assert np.allclose(np.diff(zs), h / math.sqrt(3), rtol=1e-5)
    # odd layers start at (h/sqrt 8, h/sqrt 8) and alternate rows shift by h/2 more
    odd = L[L[:, 2] == zs[1]]
    assert odd[:, 1].min() == pytest.approx(h / math.sqrt(8), rel=1e-6)
    assert odd[:, 0].min() == pytest.approx(h / math.sqrt(8), rel=1e-6)
    even = L[L[:, 2] == zs[0]]
    rows = np.unique(even[:, 1])
    assert even[even[:, 1] == rows[1]][:, 0].min() == pytest.approx(h / 2, rel=1e-6)
    # a 1-D lattice: dim 0 on the other axes gives a single row of the x step
    L1 = R.lattice(R.AABB([0], [1.0]), 0.25)
    assert len(L1) == 5 and np.all(L1[:, 1:] == 0)


def test_region_closed_forms():
    box = R.AABB([0, 0, 0], [1, 2, 3])
    p = np.array([[0, 0, 0, 0], [1, 2, 3, 0], [0.5, 1, 1.5, 0], [-1e-6, 1, 1, 0], [0.5, 2.01, 1, 0],
                  [0.5, 1, 1, 1e-3]], f32)
    assert box.contains(p).tolist() == [True, True, True, False, False, False]
    shell = R.AABB([0, 0], [1, 1]).border(0.1)
    q = np.array([[0.05, 0.5, 0, 0], [0.5, 0.5, 0, 0], [0.95, 0.95, 0, 0], [0.5, 0.09, 0, 0]], f32)
    assert shell.contains(q).tolist() == [True, False, True, True]
    e = R.AABE([0, 0], [2, 1])  # x^2/4 + y^2 <= 1, z and w radii 0 (not constrained)
    q = np.array([[1.99, 0, 0, 0], [0, 1.01, 0, 0], [math.sqrt(2), math.sqrt(0.5) * 0.999, 0, 0],
                  [0, 0, 5, 0]], f32)
    assert e.contains(q).tolist() == [True, False, True, True]
    ring = R.AABE([0, 0, 0], [1, 1, 1]).border(0.25)
    q = np.array([[0.9, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.0, 0.8, 0]], f32)
    assert ring.contains(q).tolist() == [True, False, True]
    u = R.Union(R.AABB([0, 0], [1, 1]), R.AABB([2, -1], [1, 1]))
    assert u.bound().min.tolist()[:2] == [0, -1] and u.bound().dim.tolist()[:2] == [3, 2]
    i = R.Intersection(R.AABB([0, 0], [2, 2]), R.AABB([1, 1], [2, 2]))
    assert i.bound().min.tolist()[:2] == [1, 1] and i.bound().dim.tolist()[:2] == [1, 1]
    q = np.array([[1.5, 1.5, 0, 0], [0.5, 0.5, 0, 0]], f32)
    assert i.contains(q).tolist() == [True, False]
    d = R.Difference(R.AABB([0, 0], [2, 2]), R.AABE([1, 1], [0.5, 0.5]))
    q = np.array([[1.3, 1.3, 0, 0], [0.5, 0.5, 0, 0], [1.4, 1.4, 0, 0]], f32)
    assert d.contains(q).tolist() == [False, True, True] and d.bound().dim.tolist()[:2] == [2, 2]


def test_transform_closed_forms():
    rr = R.RegionReader()
    # scale then translate: x = (p - t) / s must lie in the unit box
    r = rr.region({"min": [0.0, 0.0], "dim": [1.0, 1.0], "scale": [2.0, 3.0], "translation": [1.0, 1.0]})
    q = np.array([[2.9, 3.9, 0, 0], [3.1, 1.5, 0, 0], [1.0, 1.0, 0, 0], [0.99, 2, 0, 0]], f32)
    assert r.contains(q).tolist() == [True, False, True, False]
    b = r.bound()
    assert b.min.tolist()[:2] == [1, 1] and b.dim.tolist()[:2] == [2, 3]
    # a quarter turn about x: (x, y, z) -> (x, -z, y)
    r = rr.region({"min": [0.0, 0.0, 0.0], "dim": [1.0, 0.2, 0.1], "rotation": [1.0, 0.0, 0.0, 90.0]})
    q = np.array([[0.5, -0.05, 0.1, 0], [0.5, 0.05, 0.1, 0], [0.5, -0.05, -0.1, 0]], f32)
    assert r.contains(q).tolist() == [True, False, False]
    b = r.bound()
    assert np.allclose(b.min[:3], [0, -0.1, 0], atol=1e-7) and np.allclose(b.dim[:3], [1, 0.1, 0.2], atol=1e-7)
    # a scalar rotation turns about z
    r = rr.region({"min": [0.0, 0.0], "dim": [1.0, 1.0], "rotation": 0.0})
    assert r.contains(np.array([[0.5, 0.5, 0, 0]], f32)).tolist() == [True]


@pytest.mark.parametrize("ascii_", [False, True])
def test_mesh_and_surface_closed_forms(tmp_path, ascii_):
    path = str(tmp_path / "cube.stl")
    write_stl(path, cube_tris(), ascii_)
    m = R.read_stl(path)
    assert len(m.verts) == 8 and len(m.faces) == 12
    mesh = R.Mesh(m)
    rng = np.random.default_rng(3)
    inside = np.zeros((50, 4), f32)
    inside[:, :3] = rng.uniform(0.01, 0.99, (50, 3))
    assert mesh.contains(inside).all()
    out = np.array([[1.5, 0.5, 0.5, 0], [0.5, 0.5, -0.5, 0], [0.5, 0.5, 1.5, 0], [-0.2, 0.3, 0.3, 0],
                    [0.5, 0.5, 0.5, 1.0]], f32)
    assert not mesh.contains(out).any()
    assert mesh.bound().min.tolist() == [0, 0, 0, 0] and mesh.bound().dim.tolist() == [1, 1, 1, 0]
    s = R.Surface(m, 0.1)
    q = np.array([[0.5, 0.5, 1.05, 0], [0.5, 0.5, 0.5, 0], [1.05, 1.05, 1.05, 0], [0.5, 0.5, 1.2, 0],
                  [0.5, -0.08, 0.5, 0]], f32)
    assert s.contains(q).tolist() == [True, False, True, False, True]
    assert np.allclose(s.bound().min[:3], -0.1) and np.allclose(s.bound().dim[:3], 1.2)


# ---------------------------------------------------------------------------------------
# TOML reader
# ---------------------------------------------------------------------------------------
VALID = [
    'a = 1\nb = -2\nc = +3\nd = 1_000\ne = 0x1F\nf = 0o17\ng = 0b101\n',
    'a = 1.5\nb = -0.25\nc = 1e3\nd = 2.5E-3\ne = 6.02e+23\nf = 1_0.0_1\ng = inf\nh = -inf\n',
    'a = 0.1\nb = 0.30000000000000004\nc = 1.7976931348623157e308\nd = 5e-324\n',
    's = "tab\\there \\"q\\" \\\\ \\u00e9 \\U0001F600"\nl = \'C:\\path\'\n',
    'm = """\nline1\nline2 \\\n   joined"""\nn = \'\'\'\nraw \\n\'\'\'\n',
    'arr = [1.0, 2.0,\n  3.0, # comment\n]\nnested = [[1, 2], ["a"], []]\nempty = []\n',
    '[objects.car]\npacking = 0.5\nfloor = {contact_strength = 5e5, friction = 1.0}\n'
    '[objects.car.region]\nunion = [{region = {translation = [-0.2,-0.22], model = "x.stl"}}]\n',
    'a.b.c = 1\na.b.d = 2\n"quoted key" = true\n\'lit.key\' = false\n',
    '[[fruit]]\nname = "apple"\n[fruit.physical]\ncolor = "red"\n[[fruit]]\nname = "banana"\n',
    '# only a comment\n\n   \n',
    't = {}\nu = { a = 1, b.c = "x" }\n',
    '[x.y.z]\nw = 1\n[x]\nv = 2\n',
    'name = "Car"\r\nlighting = true\r\n',
]

INVALID = [
    'a = 1\na = 2\n',
    '[t]\nx = 1\n[t]\ny = 2\n',
    'a = "unterminated\n',
    'a = 01\n',
    'a = 1__0\n',
    'a = \n',
    'a = 1.\n',
    'a = .5\n',
    'a = [1, 2\n',
    '[a]\nb = 1\n[a.b]\nc = 2\n',
    't = {a = 1}\n[t.b]\nc = 1\n',
    'a = "bad \\q escape"\n',
    'key with space = 1\n',
    'a = 1 b = 2\n',
]


def _same_tree(a, b):
    if isinstance(a, dict):
        return isinstance(b, dict) and sorted(a) == sorted(b) and all(_same_tree(a[k], b[k]) for k in a)
    if isinstance(a, list):
        return isinstance(b, list) and len(a) == len(b) and all(_same_tree(x, y) for x, y in zip(a, b))
    if isinstance(a, float):
        return isinstance(b, float) and (a == b or (math.isnan(a) and math.isnan(b)))
    return type(a) is type(b) and a == b


@pytest.mark.parametrize("doc", VALID, ids=range(len(VALID)))
def test_toml_reader_matches_tomli(doc):
    got = K.toml_to_json(doc)
    want = tomli.loads(doc)
    assert _same_tree(got, want), (got, want)


@pytest.mark.parametrize("doc", INVALID, ids=range(len(INVALID)))
def test_toml_reader_rejects(doc):
    with pytest.raises(tomli.TOMLDecodeError):
        tomli.loads(doc)
    with pytest.raises(ArpError, match="TOML line"):
        K.toml_to_json(doc)


def test_toml_reader_refuses_what_scenes_never_hold():
    # dates (valid TOML, never a scene value) and newlines inside inline tables (refused
    # by TOML 1.0 and the toml 0.5 crate alike)
    for doc in ('a = 1979-05-27\n', 't = 07:32:00\n', 'a = {b = 1,\nc = 2}\n'):
        with pytest.raises(ArpError, match="TOML line"):
            K.toml_to_json(doc)
    # nesting is bounded (the reader recurses); 100 levels still read
    with pytest.raises(ArpError, match="nested"):
        K.toml_to_json("a = " + "[" * 200 + "]" * 200 + "\n")
    assert K.toml_to_json("a = " + "[" * 100 + "]" * 100 + "\n")["a"] is not None


# ---------------------------------------------------------------------------------------
# scenes against the oracle
# ---------------------------------------------------------------------------------------
SYNTHETIC = """\
name = "Every key"
kernel_radius = 0.05
time_step = 0.0002
gravity = 2.5
artificial_viscocity = 0.3
subticks = 3
integrator = "Heun"
min = [-1.0, -1.0, -1.0]
dim = [2.0, 2.0, 2.0]

[boundaries.floor]
  packing = 0.6
  friction = 0.5
  color = [0.1, 0.2, 0.3, 1.0]
  union = [
    {min = [-1.0, -1.0, -1.0], dim = [2.0, 0.1, 2.0]},
    {center = [0.0, -0.6, 0.0], radii = [0.3, 0.2, 0.3], border = 0.08},
  ]
  gas = {potential = "Hertz", contact_factor = 1.5, dampening = 0.25}

[boundaries.bowl]
  spacing = 0.035
  region = {center = [0.5, 0.4, 0.0], diameters = [0.4, 0.4, 0.4], border = 0.06}

[objects.gas]
  speed_of_sound = 30.0
  start_density = 1.2
  target_density = 1.0
  viscocity = 0.01
  region = {min = [-0.5, 0.0, -0.5], dim = [0.4, 0.3, 0.4], rotation = [0.0, 1.0, 0.0, 30.0]}
  velocity = [0.5, 0.0, -0.25]

[objects.jelly]
  density = 2.0
  normal_stiffness = -20000.0
  shear_stiffness = -10000.0
  normal_dampening = -5.0
  shear_dampening = -4.0
  strain_order = 1
  yield_strength = 300.0
  relaxation_time = 0.5
  velocity = [0.0, -1.0, 0.0]
  angular_velocity = [0.0, 2.0, 0.5]
  floor = {friction = 0.8, contact_strength = 5000.0}
  water = {potential = "LennardJones", contact_radius = 0.07}
  difference = [
    {min = [0.0, 0.0, 0.0], dim = [0.3, 0.3, 0.3], scale = [1.0, 1.5, 1.0], translation = [0.05, 0.0, 0.0]},
    {center = [0.15, 0.15, 0.15], radii = [0.08, 0.08, 0.08]},
  ]

[objects.water]
  speed_of_sound = 40.0
  density = 1000.0
  viscocity = 0.05
  state = "Tait"
  intersection = [
    {center = [-0.4, 0.5, 0.3], dim = [0.4, 0.4, 0.3]},
    {center = [-0.4, 0.5, 0.3], radii = [0.25, 0.25, 0.3]},
  ]

[objects.block]
  density = 3.0
  friction = 0.2
  state_equation = "constant"
  region = {model = "cube.stl", scale = 0.2, translation = [0.4, -0.3, -0.4]}

[objects.shell]
  density = 1.0
  region = {model = "cube.stl", border = 0.12, scale = [0.25, 0.25, 0.25], translation = [-0.2, -0.3, 0.4]}

[objects.spawn]
  on_click = "Relative"
  density = 1.0
  region = {center = [0.0, 0.0, 0.0], radii = [0.1, 0.1, 0.1]}
"""


@pytest.fixture()
def synthetic_dir(tmp_path):
    write_stl(str(tmp_path / "cube.stl"), cube_tris())
    return str(tmp_path)


def test_synthetic_scene_matches_oracle(synthetic_dir):
    ref = oracle_scene(SYNTHETIC, synthetic_dir)
    sc = K.loads(SYNTHETIC, base_dir=synthetic_dir)
    assert_same_scene(sc, ref)
    assert sc.name == "Every key" and sc.integrator == "heun" and sc.subticks == 3
    assert sc.h == float(f32(0.05)) and sc.g == float(f32(2.5)) and sc.alpha == float(f32(0.3))
    names = [bytes(o["name"]).rstrip(b"\0").decode() for o in sc.objects]
    # boundaries first, then objects, each in byte order; on_click objects last
    assert names == ["bowl", "floor", "block", "gas", "jelly", "shell", "water", "spawn"]
    assert [int(o["on_click"]) for o in sc.objects] == [0] * 7 + [2]
    assert sc.n > 100 and sc.nb > 100 and sc.ns > 10
    # the configured pairs, in either order, and defaults elsewhere
    nm = len(sc.materials)
    it = sc.interactions.reshape(nm, nm)
    assert it[1, 3]["potential"] == 5 and it[3, 1]["potential"] == 5  # floor <-> gas
    assert it[4, 6]["potential"] == 4 and it[4, 6]["radius"] == f32(0.07)  # jelly <-> water
    assert it[1, 4]["strength"] == f32(5000.0)
    assert it[3, 6]["potential"] == 0  # gas <-> water: two fluids, no contact
    st = K.SceneFile(SYNTHETIC, synthetic_dir)
    # every particle of an object is inside its region, per the native membership test
    for i, o in enumerate(sc.objects[:7]):
        arr = sc.boundary if o["immobile"] else sc.particles
        pts = arr["pos"][int(o["first"]):int(o["first"]) + int(o["count"])]
        assert st.contains(i, pts).all()
    st.close()


def test_generation_is_thread_count_independent(synthetic_dir):
    f = K.SceneFile(SYNTHETIC, synthetic_dir)
    a = f.generate(threads=1)
    b = f.generate(threads=7)
    assert a.particles.tobytes() == b.particles.tobytes() and a.boundary.tobytes() == b.boundary.tobytes()
    f.close()


ERRORS = [
    ('kernel_radius = 1\n', "must be a float"),
    ('subticks = 1.0\n', "must be an integer"),
    ('[objects.a]\nstate = "plasma"\nregion = {min = [0.0, 0.0], dim = [1.0, 1.0]}\n', "invalid state equation"),
    ('[objects.a]\ndensity = 1.0\n', "no table named region"),
    ('[objects.a]\nregion = {min = [0, 0], dim = [1.0, 1.0]}\n', "must be floats"),
    ('[objects.a]\nregion = {model = "missing.stl"}\n', "cannot open STL model"),
    ('a = [1.0\n', "TOML line"),
]


@pytest.mark.parametrize("text,msg", ERRORS, ids=range(len(ERRORS)))
def test_reference_errors(tmp_path, text, msg):
    with pytest.raises(ArpError, match=msg):
        K.loads(text, base_dir=str(tmp_path))


def _examples():
    if not os.path.isdir(REF_EXAMPLES):
        return []
    return sorted(f for f in os.listdir(REF_EXAMPLES) if f.endswith(".toml"))


def _has_model(doc):
    return "model" in repr(doc)


@pytest.mark.parametrize("name", _examples() or ["absent"])
def test_reference_examples(name):
    """The reference's own scenes: bit-exact with the oracle (scenes with STL models: the
    lattice, masses and counts in full, membership on a sample of lattice points)."""
    if name == "absent":
        pytest.skip("reference examples not present")
    base = os.path.dirname(REF_EXAMPLES)
    text = open(os.path.join(REF_EXAMPLES, name)).read()
    doc = tomli.loads(text)
    try:
        cfg = R.parse(doc, base)
    except FileNotFoundError as e:
        # model_import_test.toml names a model the reference does not ship: it panics there
        with pytest.raises(ArpError, match="cannot open STL model"):
            K.loads(text, base_dir=base)
        print(name, "missing model:", e)
        return
    if not _has_model(doc):
        assert_same_scene(K.loads(text, base_dir=base), R.generate(cfg))
        return
    f = K.SceneFile(text, base)
    sc = f.generate()
    rng = np.random.default_rng(0)
    for i, o in enumerate(cfg["objects"]):
        rec = sc.objects[i]
        sp = f32(o["packing"] * cfg["h"])
        L = R.lattice(o["region"].bound(), sp)
        assert int(rec["lattice_points"]) == len(L)
        assert rec["material"]["mass"] == R.lattice_mass(o["mat"]["start_den"], o["region"].bound(), len(L))
        arr = sc.boundary if o["mat"]["immobile"] else sc.particles
        got = arr["pos"][int(rec["first"]):int(rec["first"]) + int(rec["count"])]
        have = {bytes(r) for r in np.ascontiguousarray(got)}
        # the generated points are lattice points, in lattice order
        idx = np.nonzero(np.isin(L.view(np.dtype((np.void, 16))).ravel(),
                                 np.ascontiguousarray(got).view(np.dtype((np.void, 16))).ravel()))[0]
        assert len(idx) == len(got) and np.array_equal(L[idx], got)
        sample = rng.choice(len(L), size=min(len(L), 300), replace=False)
        want = o["region"].contains(L[sample])
        assert np.array_equal(f.contains(i, L[sample]), want), (name, o["name"])
        assert [bytes(L[k]) in have for k in sample] == want.tolist(), (name, o["name"])
    f.close()


def test_add_particles_oracle_rules():
    """The oracle's Particles::add_particles: equal fluid materials are reused, a new one
    gets default_between everywhere (itself included), a stiff one is always new with a
    zero self-interaction, and its particles become solids replacing the placeholder."""
    from arp import wire

    h = np.float32(0.02)
    wall = np.zeros((), wire.MATERIAL)
    wall["immobile"], wall["visc"], wall["start_den"], wall["target_den"] = 1, 0.5, 1.0, 1.0
    water = np.zeros((), wire.MATERIAL)
    water["state_eq"], water["start_den"], water["sound_speed"], water["visc"] = 3, 1000.0, 20.0, 0.1
    mats = np.array([wall, water], wire.MATERIAL)
    inter = np.zeros(4, wire.INTERACTION)
    inter[1]["potential"] = inter[2]["potential"] = 2
    parts = wire.particles(3)
    parts["mat"] = 1
    solids = wire.solids(1)
    solids["part_id"] = wire.INVALID
    new = wire.particles(2)
    new["pos"][:, 0] = [1.0, 2.0]
    # an equal fluid: reused, nothing grows
    P, S, M, I, mid, added = R.add_particles(parts, solids, mats, inter, water.copy(), new, h)
    assert (mid, added, len(M), len(I), len(P)) == (1, False, 2, 4, 5) and list(P["mat"][3:]) == [1, 1]
    # a different fluid: a new material, Linear contact (radius 2h, friction = sum of
    # viscosities) with the wall, none with water or itself
    oil = water.copy()
    oil["start_den"] = 800.0
    P, S, M, I, mid, added = R.add_particles(P, S, M, I, oil, new, h)
    assert (mid, added, len(M), len(I)) == (2, True, 3, 9)
    it = I.reshape(3, 3)
    assert it[0, 2]["potential"] == 2 and it[2, 0]["potential"] == 2
    assert it[0, 2]["radius"] == np.float32(0.04) and it[0, 2]["friction"] == np.float32(0.6)
    assert it[1, 2]["potential"] == 0 and it[2, 2]["potential"] == 0 and it[0, 1]["potential"] == 2
    # a zero-EOS material touches itself with Linear contact
    sand = water.copy()
    sand["state_eq"] = 0
    _, _, _, I2, _, _ = R.add_particles(P, S, M, I, sand, new, h)
    assert I2.reshape(4, 4)[3, 3]["potential"] == 2
    # a stiff material: new every time, zero self-interaction, solids replace the placeholder
    rub = water.copy()
    rub["normal_stiffness"] = -1e4
    P, S, M, I, mid, added = R.add_particles(P, S, M, I, rub, new, h)
    assert (mid, added) == (3, True) and I.reshape(4, 4)[3, 3]["potential"] == 0
    assert list(S["part_id"]) == [7, 8] and list(P["solid_id"][7:]) == [0, 1]
    assert np.array_equal(S["ref_pos"], new["pos"])
    P, S, M, I, mid, added = R.add_particles(P, S, M, I, rub.copy(), new, h)
    assert (mid, added) == (4, True) and list(S["part_id"]) == [7, 8, 9, 10]
