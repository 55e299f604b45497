"""Write the reference golden vectors under tests/golden/.

Runs the compiled reference driver (oracle/_ref/ref_driver, built from
/root/reference/src by `make -f oracle/Makefile.ref`) on the reference's own
built-in case and packs its raw dumps into compressed .npz fixtures:

* ref_geometry.npz  - connectivity, metrics and boundary tables of the
                      36 x 6 x 12 O-grid (create_hollow ... fill_global_arrays,
                      evaluate_all_mesh_constants, evaluate_pressure_bc_array,
                      calculate_pressure_nim_constants).
* ref_steps.npz     - the seeded initial state and the state after time steps
                      1 and NSTEPS, with per-step iteration counts and the
                      values the reference prints (nse.f90:169-171,
                      nse.f90:2792-2796).
* ref_iteration.npz - every intermediate array of one instrumented outer
                      iteration (coefficients, sources, each sweep, each
                      boundary update, find_error) starting from the state
                      after step NSTEPS.

* ref_c1.json       - the same dumps from the reference built at the 48 x 24
                      x 16 benchmark case (oracle/_ref/c1/ref_driver, `make -f
                      oracle/Makefile.ref c1`), 2 steps: for every array its
                      length, SHA-256 of its little-endian bytes and 8 sample
                      values (the arrays themselves would be about 60 MB),
                      plus the printed step lines.

Only data is stored. Re-run with `python tests/golden/make_golden.py` in a
container that has the reference checked out (`--c1` writes ref_c1.json only).
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
DRIVER_C1 = os.path.join(ROOT, "oracle", "_ref", "c1", "ref_driver")
NSTEPS_C1 = 2
NSTEPS = 3
NOISE_AMP = 1e-3
SEED = 20221


def _load(d, name):
    for ext, dt in ((".f64", "<f8"), (".i32", "<i4")):
        p = os.path.join(d, name + ext)
        if os.path.exists(p):
            return np.fromfile(p, dtype=dt)
    raise FileNotFoundError(name)


def _parse_stdout(text):
    steps = []
    pat = re.compile(r"step #\s*(\d+)\s*\|\s*# iters.\s*(\d+)\s*\|\s*err V=\s*(\S+)\s*\|\s*err p=\s*(\S+)"
                     r"\s*\|\s*err T=\s*(\S+)\s*\|\s*D=\s*(\S+)\s*\|\s*dudt=\s*(\S+)\s*\|\s*dTdt=\s*(\S+)")
    nu = re.compile(r"Nu_H=\s*(\S+)\s*\|\s*Nu_C=\s*(\S+)\s*\|\s*Average =\s*(\S+)\s*\|\s*Area1 =\s*(\S+)"
                    r"\s*\|\s*Area2 =\s*(\S+)")
    nus = []
    for line in text.splitlines():
        m = pat.search(line)
        if m:
            steps.append([float(x) for x in m.groups()])
        m = nu.search(line)
        if m:
            nus.append([float(x) for x in m.groups()])
    return np.array(steps), np.array(nus)


def digest(a):
    """Length, SHA-256 of the little-endian bytes and 8 evenly spaced samples."""
    a = np.ascontiguousarray(a)
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    idx = np.linspace(0, a.size - 1, 8).astype(np.int64) if a.size else np.zeros(0, dtype=np.int64)
    return {"n": int(a.size), "dtype": a.dtype.str, "sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "idx": idx.tolist(), "sample": a[idx].tolist()}


def write_c1():
    if not os.path.exists(DRIVER_C1):
        subprocess.check_call(["make", "-f", "oracle/Makefile.ref", "c1"], cwd=ROOT)
    with tempfile.TemporaryDirectory() as d:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        out = subprocess.run([DRIVER_C1, d, str(NSTEPS_C1), repr(NOISE_AMP), str(SEED)],
                             env=env, check=True, capture_output=True, text=True).stdout
        names = sorted({os.path.splitext(f)[0] for f in os.listdir(d)})
        arrs = {n: _load(d, n) for n in names}
    steps, nus = _parse_stdout(out)
    doc = {"case": "48 x 24 x 16 (theta x r x z), R_i 0.3, R_o 1, L 1, Ra 1e4, Pr 0.71, DT 1e-3, "
                   "blend radii 0.44899385947278925 / 0.8163312864909298",
           "nsteps": NSTEPS_C1, "noise_amp": NOISE_AMP, "seed": SEED,
           "printed_steps": steps.tolist(), "printed_nu": nus.tolist(),
           "arrays": {k: digest(v) for k, v in arrs.items() if not k.endswith("_A")}}
    path = os.path.join(HERE, "ref_c1.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=0, sort_keys=True)
    print("ref_c1.json", os.path.getsize(path))
    return 0


def main():
    if "--c1" in sys.argv:
        return write_c1()
    if not os.path.exists(DRIVER):
        subprocess.check_call(["make", "-f", "oracle/Makefile.ref"], cwd=ROOT)
    with tempfile.TemporaryDirectory() as d:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        out = subprocess.run([DRIVER, d, str(NSTEPS), repr(NOISE_AMP), str(SEED)],
                             env=env, check=True, capture_output=True, text=True).stdout
        names = sorted({os.path.splitext(f)[0] for f in os.listdir(d)})
        arrs = {n: _load(d, n) for n in names}

    geo = {k[4:]: v for k, v in arrs.items() if k.startswith("geo_")}
    np.savez_compressed(os.path.join(HERE, "ref_geometry.npz"), **geo)

    steps, nus = _parse_stdout(out)
    st = {"printed_steps": steps, "printed_nu": nus,
          "meta": np.array([NSTEPS, NOISE_AMP, SEED], dtype=np.float64)}
    for k, v in arrs.items():
        for pre in ("s0_", "s1_", "s%d_" % NSTEPS):
            if k.startswith(pre):
                st[k] = v
    np.savez_compressed(os.path.join(HERE, "ref_steps.npz"), **st)

    # the exponential-fit intermediates M_A/MT_A are pinned through the
    # coefficient arrays built from them and are not stored
    it = {k: v for k, v in arrs.items() if k.startswith("it") and not k.endswith("_A")}
    np.savez_compressed(os.path.join(HERE, "ref_iteration.npz"), **it)
    for f in ("ref_geometry.npz", "ref_steps.npz", "ref_iteration.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))
    return write_c1()


if __name__ == "__main__":
    sys.exit(main())
