"""rainbow_amd: the prox_rigid_bodies time step of RAINBOW on AMD MI355X (gfx950).

* ``rainbow_amd.api``      the reference's scene/simulate API (drop-in);
* ``rainbow_amd.stepper``  HipStepper, the GPU stepper behind ``engine.stepper``;
* ``rainbow_amd.world``    device-resident batches of scenes (batched stepping);
* ``rainbow_amd.serving``  continuous-batching serving loop;
* ``rainbow_amd.native``   ctypes binding of librbh (include/rbh.h).
"""
__version__ = "0.1.0"
