"""The Fortran drop-in driver nim3d_hip end to end on the reference's own
case: mesh from the Fortran host, metrics on the device (or the host), 3 time
steps through the C-ABI
from the seeded initial state, restart file in the reference's record layout
(restart.f90:8-39), compared with the reference's own state after step 3
(tests/golden/ref_steps.npz, written by the compiled reference).

The reference converged each step to its own tolerances (1e-6 on u, v, w,
T, 5e-3 on p), in its in-place order; the driver converges to the same
tolerances in the 4-colour order (FP64) or on the FP64 residual (mixed). The
two iterates are therefore only as close as those tolerances make them.
"""
import os
import subprocess

import numpy as np
import pytest

import gpucase as G
import refdata as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nim-conv22-shared_amd", "nim3d_hip")
NML = """&case n_theta=36, n_r=6, n_z=12, R_i=1.0471975511966667, R_o=2., height=2.,
      blend_in=1.25, blend_out=1.75, Ra=1e5, Pr=0.7, DT=0.01, nsteps=3,
      tol=1e-6,1e-6,1e-6,5e-3,1e-6, noise=1e-3, seed=20221, check_steady=.false.,
      final_out='{out}', mixed_inner={inner} /
"""
RECORDS = ["faces_u", "faces_v", "faces_w", "faces_p", "faces_T", "uxyz", "vxyz", "wxyz", "Txyz",
           "uxyz_old", "vxyz_old", "wxyz_old", "Txyz_old", "meshes_dila"]


def read_restart(path):
    """Fortran unformatted sequential records: TI, STEP, then the arrays."""
    data = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "little")
        recs.append(data[pos + 4:pos + 4 + n])
        pos += n + 8
    out = {"TI": np.frombuffer(recs[0], dtype=np.float64)[0], "STEP": int.from_bytes(recs[1], "little")}
    for name, r in zip(RECORDS, recs[2:]):
        out[name] = np.frombuffer(r, dtype=np.float64)
    return out


@pytest.fixture(scope="module")
def oracle_colored():
    """The oracle's 3 steps in the driver's (4-colour) order: iteration
    counts and final faces. In this order the oracle is 1.8e-3 (u, v),
    9e-4 (w), 2.6e-6 (T) and 3.7e-3 (p) from the reference's in-place
    order after 3 steps at these tolerances."""
    import oracle as O
    geo, st = R.geometry(), R.steps()
    o = G.oracle_context(geo)
    R.load_state(o, st, "s0")
    iters = []
    for l in range(3):
        rc, scal, nu = o.time_step(TOL, O.COLORED, min_iter=2 if l == 0 else 0)
        assert rc == 0
        iters.append(int(scal[0]))
    out = {n: o.get(n).copy() for n in ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]}
    o.close()
    return iters, out


TOL = np.array([1e-6, 1e-6, 1e-6, 5e-3, 1e-6])
# reference order against the 4-colour order after 3 steps at these
# tolerances, about twice the oracle's own gap
ORDER_BOUND = dict(uvw=5e-3, T=1e-5, p=1e-2)


def run_driver(tmp_path, inner):
    assert os.path.exists(EXE), "build the driver first (make -C nim-conv22-shared_amd/fortran)"
    out = tmp_path / "final"
    nml = tmp_path / "case.nml"
    nml.write_text(NML.format(out=out, inner=inner))
    p = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "completed 3 steps" in p.stdout, p.stdout[-2000:]
    iters = [int(l.split("# iters.")[1].split("|")[0]) for l in p.stdout.splitlines() if l.startswith("step #")]
    return read_restart(out), iters


@pytest.mark.parametrize("inner", [0, 4])
def test_driver_three_steps(tmp_path, oracle_colored, inner):
    st, iters = run_driver(tmp_path, inner)
    ref = R.steps()
    assert st["STEP"] == 4  # solve_nse starts at STEP = 1 (nse.f90:32) and adds one per step (nse.f90:1295)
    # TI starts at DT (nse.f90:31) and one step advances 2 DT (nse.f90:1296)
    np.testing.assert_allclose(st["TI"], 0.01 + 3 * 2 * 0.01, rtol=1e-14)
    if inner == 0:
        # FP64: the same passes and the same fields as the oracle in this order
        it_o, x_o = oracle_colored
        assert iters == it_o, (iters, it_o)
        for n, v in x_o.items():
            assert G.rel_l2(st[n], v) < 1e-9, n
    # both loops against the reference's own state after step 3
    diff = {}
    for f in "uvwT":
        diff["faces_" + f] = G.rel_l2(st["faces_" + f], ref["s3_faces_" + f])
        diff[f + "xyz"] = G.rel_l2(st[f + "xyz"], ref["s3_%sxyz" % f])
    pg, pr = st["faces_p"], ref["s3_faces_p"]
    diff["p"] = G.rel_l2(pg - pg.mean(), pr - pr.mean())
    for k, v in diff.items():
        bound = ORDER_BOUND["T" if "T" in k else "p" if k == "p" else "uvw"]
        assert v < bound, diff


NML_RESUME = """&case n_theta=36, n_r=6, n_z=12, R_i=1.0471975511966667, R_o=2., height=2.,
      blend_in=1.25, blend_out=1.75, Ra=1e5, Pr=0.7, DT=0.01, nsteps={nsteps},
      tol=1e-6,1e-6,1e-6,5e-3,1e-6, noise=1e-3, seed=20221, check_steady=.false.,
      restart_every={every}, restart_prefix='{prefix}', restart_in='{rin}', final_out='{out}',
      mixed_inner={inner} /
"""


@pytest.mark.parametrize("inner", [0, 4])
def test_restart_resume_is_bitwise(tmp_path, inner):
    """4 steps in one run against 2 steps, a restart file (restart.f90:8-39,
    written after the step's solve as solve_nse does) and 2 more steps from
    it (read_restart + update_time_step, nse.f90:104-108): the final states
    are bit-identical."""
    def run(name, nsteps, every=0, rin=""):
        nml = tmp_path / (name + ".nml")
        out = tmp_path / (name + ".final")
        nml.write_text(NML_RESUME.format(nsteps=nsteps, every=every, prefix=tmp_path / "rest", rin=rin, out=out,
                                         inner=inner))
        p = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        return read_restart(out)
    straight = run("straight", 4)
    run("first", 2, every=2)
    assert (tmp_path / "rest2").exists()
    resumed = run("resumed", 2, rin=tmp_path / "rest2")
    assert straight["STEP"] == resumed["STEP"] == 5
    assert straight["TI"] == resumed["TI"]
    for n in RECORDS:
        assert np.array_equal(straight[n], resumed[n]), n


def test_device_and_host_metrics_run_the_same_bits(tmp_path):
    """device_metrics=.true. (the default: nim_device_metrics) and .false.
    (the Fortran host's build_geometry) give the same final state bit for
    bit: the two metric paths produce the same arrays."""
    def run(flag):
        nml = tmp_path / ("m%s.nml" % flag)
        out = tmp_path / ("m%s.final" % flag)
        nml.write_text(NML.format(out=out, inner=0).replace(" /\n", ", device_metrics=.%s. /\n" % flag))
        p = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
        return read_restart(out)
    a, b = run("true"), run("false")
    for n in RECORDS:
        assert np.array_equal(a[n], b[n]), n


def read_vtk(path):
    """Points, hexahedra and cell scalars of the driver's legacy VTK file."""
    lines = open(path).read().split("\n")
    assert lines[0].startswith("# vtk DataFile") and lines[3] == "DATASET UNSTRUCTURED_GRID"
    n_pts = int(lines[4].split()[1])
    pts = np.array([list(map(float, l.split())) for l in lines[5:5 + n_pts]])
    c = 5 + n_pts
    n_cells = int(lines[c].split()[1])
    conn = np.array([list(map(int, l.split())) for l in lines[c + 1:c + 1 + n_cells]])
    fields, q = {}, c + 1 + n_cells
    while q < len(lines):
        if lines[q].startswith("SCALARS"):
            name = lines[q].split()[1]
            fields[name] = np.array(lines[q + 2:q + 2 + n_cells], dtype=float)
            q += 2 + n_cells
        else:
            q += 1
    return lines[1], pts, conn, fields


def test_field_output_every_step_out(tmp_path):
    """write_data every STEP_OUT steps (nse.f90:59, postprocess.f90:2-124):
    with step_out=2 over 3 steps the driver writes plots/step00002.vtk (the
    mesh and the single-precision cell fields u, v, w, p, T, D of the Silo
    file, here legacy VTK) after step 2's inner solve, and lists it in
    plots.visit. Its fields are those of the same two steps run through the
    C-ABI from Python, to single precision."""
    import nim_amd
    from nim_amd import host
    nml = tmp_path / "out.nml"
    nml.write_text(NML.format(out=tmp_path / "final", inner=0).replace(" /\n", ", step_out=2 /\n"))
    p = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert (tmp_path / "plots.visit").read_text().split() == ["plots/step00002.vtk"]
    assert not (tmp_path / "plots" / "step00001.vtk").exists()
    title, pts, conn, fields = read_vtk(tmp_path / "plots" / "step00002.vtk")
    assert "cycle 2" in title
    case = host.reference_case()
    geo = R.geometry()
    xyz = host.build_geometry(case)["node_xyz"].reshape(-1, 3)
    np.testing.assert_allclose(pts, xyz, rtol=1e-7, atol=1e-7)
    assert conn.shape == (case.ncells, 9)
    np.testing.assert_allclose(xyz[conn[:, 1:]].mean(axis=1), host.cell_centroids(host.build_geometry(case)),
                               atol=1e-12)
    # the same two steps through the C-ABI
    ctx = G.gpu_context(geo)
    ctx.upload_state(host.initial_state(case, amp=1e-3, seed=20221))
    for l in range(2):
        ctx.build_momentum_coeffs()
        ctx.inner_solve(TOL, max_iter=200000, min_iter=2 if l == 0 else 0)
        ctx.end_step()
        if l == 0:
            ctx.advance_time()
    want = {"u": ctx.get("uxyzt"), "v": ctx.get("vxyzt"), "w": ctx.get("wxyzt"), "p": ctx.get("meshes_p"),
            "T": ctx.get("Txyzt"), "D": ctx.get("meshes_dila")}
    ctx.close()
    assert sorted(fields) == sorted(want)
    for n, v in want.items():
        np.testing.assert_allclose(fields[n], v.astype(np.float32), rtol=1e-6, atol=1e-6 * np.abs(v).max(), err_msg=n)


# R_i left at the driver's default, the reference's 1*(3.14159265359/3),
# as host.reference_case() has it
NML_ACCEL = """&case n_theta=36, n_r=6, n_z=12, R_o=2., height=2.,
      blend_in=1.25, blend_out=1.75, Ra=1e5, Pr=0.7, DT=0.01, nsteps=3,
      tol=1e-8,1e-8,1e-8,1e-8,1e-8, noise=1e-3, seed=20221, check_steady=.false.,
      final_out='{out}', mixed_inner={inner}, p_measure=1, accel={depth}, accel_period={period},
      accel_scope={scope} /
"""


@pytest.mark.parametrize("inner,scope", [(0, 1), (16, 1), (0, 0)])
def test_driver_accelerated_steps_match_the_library(tmp_path, inner, scope):
    """The driver with accel > 0 (every field at 1e-8, p by its half range),
    extrapolating the pressure faces (accel_scope=1) or the whole iterate
    (0), runs the same accelerated solves as the library called from Python
    on the same mesh, metrics and seeded state: the same passes per step and
    the same bits."""
    import nim_amd
    from nim_amd import host
    depth, period = ((12, 1) if scope else (16, 8)) if inner == 0 else (10, 1)
    out = tmp_path / "final"
    nml = tmp_path / "case.nml"
    nml.write_text(NML_ACCEL.format(out=out, inner=inner, depth=depth, period=period, scope=scope))
    p = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    iters = [int(l.split("# iters.")[1].split("|")[0]) for l in p.stdout.splitlines() if l.startswith("step #")]
    st = read_restart(out)
    case = host.reference_case()
    geo = host.build_geometry(case, device=True)
    ctx = G.gpu_for_case(case, geo, host.initial_state(case, amp=1e-3, seed=20221))
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    ctx.accel_options(depth=depth, period=period)
    ctx.accel_scope(scope)
    tol = np.full(5, 1e-8)
    py_iters = []
    for l in range(3):
        ctx.build_momentum_coeffs()
        rc, n, err = ctx.inner_solve_accel(tol, inner=inner, max_iter=200000 // max(inner, 1) + 1,
                                           min_iter=2 if (l == 0 and inner == 0) else 0)
        py_iters.append((n - 1) * inner if inner else n)
        ctx.end_step()
        ctx.advance_time()
    print("driver passes", iters, "library passes", py_iters)
    assert iters == py_iters
    for n in ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]:
        np.testing.assert_array_equal(st[n], ctx.get(n), err_msg=n)
    ctx.close()
