"""Scene files: the layout of scenes/write_xml.py:12-121 and the reader.

* Layout: the element tree and attribute names of the reference's writer,
  numbers printed by numpy.array2string / str() as there, the whitespace of
  its last ET.indent pass, the interaction key printed as the tuple it is.
* Lossless files restore an engine whose bodies, shapes, forces, materials,
  hinges and parameters equal the original's, whose re-written file is the
  same text, and whose next oracle step is bitwise the original's.
"""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle.scene as OS
import oracle.stepper as OST
import rainbow_amd.api as API
from rainbow_amd import math3 as Q
from rainbow_amd import mesh as MESH
from rainbow_amd import scenes, xml_io


def sample_engine(bank, hinge=True):
    eng = scenes.new_engine(bank, dict(max_iterations=20, use_pre_stabilization=True), backend=None)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    API.create_damping_force(eng, "air", 0.01, 0.02)
    API.create_surfaces_interaction(eng, "default", "rubber", 0.3, np.array([0.9, 0.8, 0.7]))
    scenes.add_ground(eng, bank)
    box = bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    ball = bank.get("ball_8", lambda: MESH.create_sphere(0.5, 8, 8))
    scenes.add_body(eng, "a", box, np.array([-0.51, 0.52, 0.0]), q=Q.identity(), v=[0.1, 0.0, 0.0])
    scenes.add_body(eng, "b", box, np.array([0.51, 0.53, 0.0]), q=Q.Ry(0.05), w=[0.0, 0.3, 0.0])
    scenes.add_body(eng, "c", ball, np.array([0.0, 1.8, 0.3]), q=Q.Rx(0.3))
    API.connect_force(eng, "c", "air")
    API.set_body_material(eng, "c", "rubber")
    if hinge:
        API.create_hinge(eng, "ab")
        API.set_hinge(eng, "ab", "a", "b", np.array([0.0, 1.02, 0.0]), np.array([0.0, 0.0, 1.0]))
    return eng


def test_layout_of_the_reference_writer(oracle_bank32, tmp_path):
    eng = sample_engine(oracle_bank32)
    fn = str(tmp_path / "scene.xml")
    xml_io.write_xml(eng, fn)
    text = open(fn, encoding="utf-8").read()
    assert text.startswith("<?xml version='1.0' encoding='utf-8'?>\n<scene>")
    root = ET.parse(fn).getroot()
    assert [c.tag for c in root] == ["bodies", "forces", "shapes", "materials", "hinges", "params"]
    b = root.find("bodies")[3]
    assert set(b.attrib) == {"name", "idx", "shape", "pos", "rot", "vel", "spin", "mass", "inertia", "is_fixed",
                             "material"}
    body = eng.bodies["c"]
    assert b.get("pos") == np.array2string(body.r) and b.get("rot") == np.array2string(body.q)
    assert b.get("mass") == str(body.mass) and b.get("is_fixed") == "False" and b.get("material") == "rubber"
    assert [a.get("name") for a in b.findall("apply")] == ["gravity", "air"]
    g, d = root.find("forces")
    assert (g.tag, d.tag) == ("gravity", "damping")
    assert set(g.attrib) == {"name", "g", "up"} and set(d.attrib) == {"name", "alpha", "beta"}
    sh = root.find("shapes")[0]
    assert [c.tag for c in sh] == ["vertices", "triangles"]
    assert sh.find("vertices").get("count") == str(len(eng.shapes[sh.get("name")].mesh.V))
    assert set(sh.find("vertices")[0].attrib) == {"idx", "p"} and set(sh.find("triangles")[0].attrib) == {"idx", "labels"}
    it = root.find("materials")
    assert [i.get("materials") for i in it] == ["('default', 'default')", "('default', 'rubber')"]
    assert set(it[0].attrib) == {"materials", "friction", "restitution"}
    h = root.find("hinges")[0]
    assert set(h.attrib) == {"name", "idx", "parent", "parent-idx", "child", "child-idx", "origin", "axis", "mode"}
    assert np.allclose(xml_io.parse_array(h.get("origin")), [0.0, 1.02, 0.0], atol=1e-8)
    assert np.allclose(xml_io.parse_array(h.get("axis")), [0.0, 0.0, 1.0], atol=1e-8)
    names = [p.get("name") for p in root.find("params")]
    assert names == sorted(names) and "time_step" in names and "max_iterations" in names
    p = {x.get("name"): x for x in root.find("params")}
    assert (p["max_iterations"].get("value"), p["max_iterations"].get("type")) == ("20", "int")
    assert (p["use_pre_stabilization"].get("value"), p["use_pre_stabilization"].get("type")) == ("True", "bool")
    # the last of the reference's ET.indent passes (initial level 3) sets the whitespace
    lines = text.splitlines()
    assert lines[1] == "<scene>" and lines[2] == "        <bodies>" and lines[3].startswith("          <body ")
    assert lines[-1] == "      </scene>"


def fields(eng):
    out = {}
    for n, b in eng.bodies.items():
        out[n] = (b.idx, b.shape.name, b.r, b.q, b.v, b.w, b.mass, b.inertia, b.is_fixed, b.material,
                  [f.name for f in b.forces])
    return out


def test_lossless_round_trip(oracle_bank32, tmp_path):
    eng = sample_engine(oracle_bank32)
    eng.step_index = 7
    fn = str(tmp_path / "a.xml")
    xml_io.write_xml(eng, fn, lossless=True)
    back = xml_io.read_xml(fn, engine=scenes.new_engine(oracle_bank32, backend=None))
    assert back.step_index == 7
    fa, fb = fields(eng), fields(back)
    assert list(fa) == list(fb)
    for n in fa:
        for x, y in zip(fa[n], fb[n]):
            if isinstance(x, np.ndarray):
                assert np.array_equal(x, y), n
            else:
                assert x == y, n
    for n, s in eng.shapes.items():
        assert np.array_equal(s.mesh.V, back.shapes[n].mesh.V) and np.array_equal(s.mesh.T, back.shapes[n].mesh.T)
    for k, beh in eng.surfaces_interactions.storage.items():
        other = back.surfaces_interactions.storage[k]
        assert np.array_equal(np.asarray(beh.mu, float), other.mu) and beh.epsilon == other.epsilon
    assert {n: (f.force_type, vars(f)) for n, f in eng.forces.items()}.keys() == back.forces.keys()
    assert back.forces["air"].alpha == 0.01 and back.forces["air"].beta == 0.02
    for name in dir(eng.params):
        if not name.startswith("__"):
            assert getattr(eng.params, name) == getattr(back.params, name), name
    h, hb = eng.hinges["ab"], back.hinges["ab"]
    assert hb.parent.name == "a" and hb.child.name == "b"
    for x, y in ((h.arm_p, hb.arm_p), (h.arm_c, hb.arm_c), (h.axis_p, hb.axis_p), (h.axis_c, hb.axis_c)):
        assert np.allclose(x, y, atol=1e-14)


def test_rewrite_is_the_same_text(oracle_bank32, tmp_path):
    eng = sample_engine(oracle_bank32, hinge=False)
    a, b = str(tmp_path / "a.xml"), str(tmp_path / "b.xml")
    xml_io.write_xml(eng, a, lossless=True)
    xml_io.write_xml(xml_io.read_xml(a, engine=scenes.new_engine(oracle_bank32, backend=None)), b, lossless=True)
    assert open(a).read() == open(b).read()


def test_restored_scene_steps_bitwise(oracle_bank32, tmp_path):
    eng = sample_engine(oracle_bank32)
    fn = str(tmp_path / "s.xml")
    xml_io.write_xml(eng, fn, lossless=True)
    back = xml_io.read_xml(fn, engine=scenes.new_engine(oracle_bank32, backend=None))
    s1, s2 = OS.from_engine(eng), OS.from_engine(back)
    for _ in range(2):
        o1, o2 = OST.step(s1, 0.01), OST.step(s2, 0.01)
        assert o1["contacts"].K == o2["contacts"].K > 0
    for x, y in ((s1.r, s2.r), (s1.q, s2.q), (s1.v, s2.v), (s1.w, s2.w)):
        assert np.array_equal(x, y)


def test_reader_errors(tmp_path):
    fn = str(tmp_path / "bad.xml")
    open(fn, "w").write("<?xml version='1.0'?><world/>")
    with pytest.raises(ValueError):
        xml_io.read_xml(fn)
    assert np.array_equal(xml_io.parse_array("[ 1.   2.5 -3. ]"), [1.0, 2.5, -3.0])
    assert np.array_equal(xml_io.parse_array("[0 1 2]", np.int64), [0, 1, 2])
