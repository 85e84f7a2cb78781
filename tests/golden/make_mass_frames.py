"""Write tests/golden/mass_frames.json: the reference's meshes, mass properties
and body frames for the shapes whose principal moments are tied.

For a cube or a sphere the inertia tensor has (nearly) equal eigenvalues, so
the body frame chosen by mass.py:xform_model_2_body_space (the eigenvectors
of np.linalg.eig, ordered by decreasing moment) depends on the last bits of
the tensor; no reference test pins it (unit_tests/test_mass.py checks the
moments of a 4 x 5 x 6 box only).  This script imports the reference's own
surface_mesh.py and mass.py (pure numpy; run it in a container that holds the
reference at /root/reference) and records, per shape, the vertex and
triangle tables, the model-space mass properties and the body frame
(r_bf2mf, q_bf2mf, mass, I_body).  The fixture is data only.

    python tests/golden/make_mass_frames.py
"""
import json
import os
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mass_frames.json")

SHAPES = [
    ("create_box", [1.0, 1.0, 1.0]),
    ("create_box", [0.5, 0.5, 0.5]),
    ("create_box", [2.0, 2.0, 2.0]),
    ("create_box", [4.0, 5.0, 6.0]),
    ("create_box", [100.0, 1.0, 100.0]),
    ("create_sphere", [0.5, 16, 16]),
    ("create_sphere", [1.0, 8, 8]),
    ("create_sphere", [0.25, 12, 10]),
]


def main():
    sys.path.insert(0, REF)
    import rainbow.geometry.surface_mesh as MESH
    import rainbow.simulators.prox_rigid_bodies.mass as MASS
    out = []
    for fn, args in SHAPES:
        V, T = getattr(MESH, fn)(*args)
        p = MASS.compute_mass_properties(V, T, 1.0)
        r, q, m, I = MASS.xform_model_2_body_space(p)
        out.append({
            "shape": fn, "args": args,
            "V": [[float(x) for x in row] for row in V], "T": [[int(x) for x in row] for row in T],
            "mass_properties": {k: float(getattr(p, k)) for k in
                                ("mass", "x", "y", "z", "Ixx", "Iyy", "Izz", "Ixy", "Ixz", "Iyz")},
            "r_bf2mf": [float(x) for x in r], "q_bf2mf": [float(x) for x in q],
            "mass": float(m), "I_body": [float(x) for x in I],
        })
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", OUT, len(out), "shapes")


if __name__ == "__main__":
    main()
