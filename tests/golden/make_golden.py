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

Only data is stored. Re-run with `python tests/golden/make_golden.py` in a
container that has the reference checked out.
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
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


def main():
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
    return 0


if __name__ == "__main__":
    sys.exit(main())
