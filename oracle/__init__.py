"""FP64 CPU oracle for the prox_rigid_bodies time step.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (``rainbow_amd``) may
import, call or link anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and there only as the checker (or as the timed CPU baseline), never as the thing
measured or shipped.

The oracle restates, in NumPy FP64, the algorithm of the reference simulator
``rainbow/simulators/prox_rigid_bodies`` (diku-dk/RAINBOW): state gather and
forces (state_storage.py), world-space triangle AABBs and the candidate-triangle
set (collision_detection.py:10-84, kdop.py, kdop_bvh.py), the Frank-Wolfe
narrow phase against trilinear SDF grids (collision_detection.py:87-201,
grid3.py:107-300), contact reduction (collision_detection.py:255-290), contact
problem assembly (constraints/contacts.py:31-287), the proximal Gauss-Seidel
solve with adaptive r-factors (gauss_seidel_solver.py:14-101,
prox_operators.py:14-34) and the semi-implicit integration
(state_storage.py:239-302).  It also restates the anisotropic ellipsoid prox
(proximal_contact/proximal_operators.py:42-139), the Jacobi sweeps
(proximal_contact/jacobi_solver.py) and the Vivace random graph colouring
(proximal_contact/blocking.py:62-111) in the deterministic form this build
defines (see ``oracle/coloring.py``).

Each module header lists how its functions are pinned (reference unit tests,
closed forms, brute force); a function with no such check says
"parity unpinned".
"""
