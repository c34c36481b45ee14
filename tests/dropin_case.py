"""Shared set-up of the drop-in tests: the reference's unchanged driver (src/main.cpp)
linked against libtsr.so, run on the reference's own test meshes and Keff table with an
options file in its format (src/main.cpp:45-146)."""
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
DRIVERS = {"libtsr loaders": os.path.join(ROOT, "oracle", "_ref", "tsr_refdriver"),
           "reference loaders": os.path.join(ROOT, "oracle", "_ref", "tsr_refdriver_refio")}

# problem_options.txt in the order main.cpp reads it, with the 18th value (the boundary
# flux) that the shipped file lacks (main.cpp:143-144): meshes, solution file, h0 [m],
# t_end [h], stored solutions and their times, soil A and soil B (Ks [m/h], theta_r,
# theta_s, alpha [1/m], n), boundary flux [m/h]
H0, TEND, TSPAN, BFLUX = -2.0, 2.0, (1.0, 2.0), 0.001
SOIL_A = (0.044, 0.058, 0.41, 7.3, 1.89)
SOIL_B = (0.044e-3, 0.058, 0.41, 7.3, 1.89)


def write_case(d):
    for f in ("macro_structured_10by10.msh", "microB_unstructured_10by10.msh", "Keff.txt"):
        shutil.copy(os.path.join(GOLDEN, f), d)
    vals = ["macro_structured_10by10.msh", "microB_unstructured_10by10.msh", "solution.txt", repr(H0), repr(TEND),
            str(len(TSPAN)), " ".join(repr(t) for t in TSPAN)]
    vals += [repr(x) for x in SOIL_A + SOIL_B] + [repr(BFLUX)]
    with open(os.path.join(d, "problem_options.txt"), "w") as f:
        f.write("\n".join(vals) + "\n")


def run_driver(exe, d, parallel, timeout=300):
    env = dict(os.environ, parallel=str(parallel), OMP_NUM_THREADS="1")
    return subprocess.run([exe, "problem_options.txt"], cwd=d, env=env, capture_output=True, text=True,
                          timeout=timeout)


def read_solution(path):
    with open(path) as f:
        n = int(f.readline())
        rows = [np.array(f.readline().split(), dtype=np.float64) for _ in range(n)]
    return np.array([r[0] for r in rows]), np.array([r[1:] for r in rows])


def read_statistics(path):
    a = np.loadtxt(path, ndmin=2)
    return a[:, 0].astype(int), a[:, 1], a[:, 2], a[:, 3].astype(int)
