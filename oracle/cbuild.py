"""Build the oracle's C helpers (oracle/csrc/*.c) into oracle/_build/ with gcc."""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")


def build():
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.c")))
    if not srcs:
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    out = os.path.join(OUT_DIR, "liboracle.so")
    tmp = out + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", tmp] + srcs + ["-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out
