"""Scene files in the reference's XML layout, and a reader for them.

``write_xml`` writes the document of scenes/write_xml.py:12-121: a <scene>
root with <bodies> (name, idx, shape, pos, rot, vel, spin, mass, inertia,
is_fixed, material and one <apply name=...> per connected force), <forces>
(<gravity name g up>, <damping name alpha beta>), <shapes> (body-frame
vertices <v idx p> and triangles <t idx labels>), <materials> (<interaction
materials friction restitution>), <hinges> (name, idx, parent, parent-idx,
child, child-idx, world origin and axis, mode="world") and <params> (every
Parameters attribute as <param name value type>, in dir() order).  Arrays
are printed by numpy.array2string and scalars by str(), as there; with
``lossless=True`` arrays use the shortest round-trip digits instead and the
root carries the engine's step count (``step_index``, which keys the
colouring's random draws), so that ``read_xml`` restores the state bit for
bit (the default keeps numpy's eight digits, like the reference).

``read_xml`` builds an Engine from such a file through the public api: the
parameters first (resolution and envelope size the SDF grids), then
materials, forces, shapes (meshes already in body frames, so no second
body-frame transform), bodies with their stored mass properties and state,
and hinges (set_hinge in world mode from the stored joint frames).
"""
import ast
import xml.etree.ElementTree as ET

import numpy as np

from . import api as API
from . import math3 as Q
from . import mesh as MESH


def _arr(a, lossless):
    a = np.asarray(a)
    if lossless and a.dtype.kind == "f":
        return np.array2string(a, floatmode="unique", max_line_width=1 << 30, threshold=1 << 30)
    return np.array2string(a)


def _num(x, lossless):
    return repr(float(x)) if lossless and isinstance(x, (float, np.floating)) else str(x)


def write_xml(engine, xml_filename: str, lossless: bool = False) -> None:
    """scenes/write_xml.py:12-121."""
    root = ET.Element("scene")
    bodies = ET.SubElement(root, "bodies")
    for body in engine.bodies.values():
        e = ET.SubElement(bodies, "body")
        e.set("name", body.name)
        e.set("idx", str(body.idx))
        e.set("shape", body.shape.name)
        e.set("pos", _arr(body.r, lossless))
        e.set("rot", _arr(body.q, lossless))
        e.set("vel", _arr(body.v, lossless))
        e.set("spin", _arr(body.w, lossless))
        e.set("mass", _num(body.mass, lossless))
        e.set("inertia", _arr(body.inertia, lossless))
        e.set("is_fixed", str(body.is_fixed))
        e.set("material", body.material)
        for force in body.forces:
            ET.SubElement(e, "apply").set("name", force.name)
    forces = ET.SubElement(root, "forces")
    for force in engine.forces.values():
        if force.force_type == "Gravity":
            e = ET.SubElement(forces, "gravity")
            e.set("name", force.name)
            e.set("g", _num(force.g, lossless))
            e.set("up", _arr(force.up, lossless))
        if force.force_type == "Damping":
            e = ET.SubElement(forces, "damping")
            e.set("name", force.name)
            e.set("alpha", _num(force.alpha, lossless))
            e.set("beta", _num(force.beta, lossless))
    shapes = ET.SubElement(root, "shapes")
    for shape in engine.shapes.values():
        e = ET.SubElement(shapes, "shape")
        e.set("name", shape.name)
        vs = ET.SubElement(e, "vertices")
        vs.set("count", str(len(shape.mesh.V)))
        for i, v in enumerate(shape.mesh.V):
            n = ET.SubElement(vs, "v")
            n.set("idx", str(i))
            n.set("p", _arr(v, lossless))
        ts = ET.SubElement(e, "triangles")
        ts.set("count", str(len(shape.mesh.T)))
        for i, t in enumerate(shape.mesh.T):
            n = ET.SubElement(ts, "t")
            n.set("idx", str(i))
            n.set("labels", _arr(t, lossless))
    materials = ET.SubElement(root, "materials")
    for key, inter in engine.surfaces_interactions.storage.items():
        e = ET.SubElement(materials, "interaction")
        e.set("materials", str(key))
        e.set("friction", _arr(inter.mu, lossless))
        e.set("restitution", _num(inter.epsilon, lossless))
    hinges = ET.SubElement(root, "hinges")
    for idx, h in enumerate(engine.hinges.values()):
        e = ET.SubElement(hinges, "hinge")
        e.set("name", h.name)
        e.set("idx", str(idx))
        e.set("parent", h.parent.name)
        e.set("parent-idx", str(h.parent.idx))
        e.set("child", h.child.name)
        e.set("child-idx", str(h.child.idx))
        e.set("origin", _arr(Q.rotate(h.parent.q, h.arm_p) + h.parent.r, lossless))
        e.set("axis", _arr(Q.rotate(h.parent.q, h.axis_p), lossless))
        e.set("mode", "world")
    params = ET.SubElement(root, "params")
    for name in dir(engine.params):
        if name.startswith("__"):
            continue
        value = getattr(engine.params, name)
        e = ET.SubElement(params, "param")
        e.set("name", str(name))
        e.set("value", _num(value, lossless) if isinstance(value, float) else str(value))
        e.set("type", type(value).__name__)
    if lossless:
        root.set("step_index", str(int(getattr(engine, "step_index", 0))))
    tree = ET.ElementTree(root)
    # write_xml.py:108-112 indents four times; each pass rewrites all the
    # whitespace, so the layout is that of the last one (initial level 3)
    ET.indent(tree, "  ", level=3)
    tree.write(xml_filename, encoding="utf-8", xml_declaration=True, method="xml", short_empty_elements=True)


def parse_array(text, dtype=np.float64):
    """numpy.array2string output ("[ 1.  2.5 -3. ]", possibly wrapped) -> array."""
    body = text.strip()
    if body.startswith("[") and body.endswith("]"):
        body = body[1:-1]
    return np.array([float(t) for t in body.replace(",", " ").split()], dtype=np.float64).astype(dtype)


_PARAM_TYPES = {"float": float, "int": int, "str": str, "bool": lambda s: s == "True"}


def read_xml(xml_filename: str, engine=None):
    """An Engine holding the scene of ``xml_filename`` (a new one, backend
    "hip", unless ``engine`` is given: its sdf_builder and stepper are kept)."""
    root = ET.parse(xml_filename).getroot()
    if root.tag != "scene":
        raise ValueError("read_xml(): not a scene file: root element <%s>" % root.tag)
    eng = engine if engine is not None else API.create_engine()
    for p in root.find("params") if root.find("params") is not None else []:
        conv = _PARAM_TYPES.get(p.get("type"))
        if conv is not None and hasattr(eng.params, p.get("name")):
            setattr(eng.params, p.get("name"), conv(p.get("value")))
    for it in root.find("materials") if root.find("materials") is not None else []:
        key = tuple(ast.literal_eval(it.get("materials")))
        mu = parse_array(it.get("friction"))
        eps = float(it.get("restitution"))
        if key in eng.surfaces_interactions.storage:
            beh = eng.surfaces_interactions.storage[key]
            beh.mu, beh.epsilon = mu, eps
        else:
            API.create_surfaces_interaction(eng, key[0], key[1], eps, mu)
    for f in root.find("forces") if root.find("forces") is not None else []:
        if f.tag == "gravity":
            API.create_gravity_force(eng, f.get("name"), float(f.get("g")), parse_array(f.get("up")))
        elif f.tag == "damping":
            API.create_damping_force(eng, f.get("name"), float(f.get("alpha")), float(f.get("beta")))
    for s in root.find("shapes") if root.find("shapes") is not None else []:
        vs = sorted(s.find("vertices"), key=lambda n: int(n.get("idx")))
        ts = sorted(s.find("triangles"), key=lambda n: int(n.get("idx")))
        V = np.array([parse_array(v.get("p")) for v in vs]).reshape(-1, 3)
        T = np.array([parse_array(t.get("labels"), np.int64) for t in ts]).reshape(-1, 3)
        API.create_shape(eng, s.get("name"), MESH.Mesh(V, T), transform_to_body_frame=False)
    for b in sorted(root.find("bodies"), key=lambda n: int(n.get("idx"))):
        name = b.get("name")
        API.create_rigid_body(eng, name)
        API.connect_shape(eng, name, b.get("shape"))
        body = eng.bodies[name]
        body.r = parse_array(b.get("pos"))
        body.q = parse_array(b.get("rot"))
        body.v = parse_array(b.get("vel"))
        body.w = parse_array(b.get("spin"))
        body.mass = float(b.get("mass"))
        body.inertia = parse_array(b.get("inertia"))
        body.is_fixed = b.get("is_fixed") == "True"
        body.material = b.get("material")
        body.idx = int(b.get("idx"))
        for a in b.findall("apply"):
            API.connect_force(eng, name, a.get("name"))
    for h in sorted(root.find("hinges") if root.find("hinges") is not None else [], key=lambda n: int(n.get("idx"))):
        API.create_hinge(eng, h.get("name"))
        API.set_hinge(eng, h.get("name"), h.get("parent"), h.get("child"), parse_array(h.get("origin")),
                      parse_array(h.get("axis")), mode=h.get("mode", "world"))
    eng.step_index = int(root.get("step_index", eng.step_index if hasattr(eng, "step_index") else 0))
    API._touch(eng)
    return eng
