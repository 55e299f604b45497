"""Order of accuracy against a closed form (the find_rms harness,
nse.f90:2801-2839). With no buoyancy (g_beta = 0 in the momentum sources)
and no initial flow the velocity stays zero and the energy equation is pure
conduction. On an undeformed annulus (blend radii inside the hexagon radius,
so no node moves) with the hexagon wall at T = 1, the cylinder at T = 0 and
adiabatic top and bottom, the steady temperature is
T(r) = ln(r / R_o) / ln(R_i / R_o) and the inner-wall Nusselt number
(scaled by R_i, as find_Nu2 prints it, nse.f90:2794) is 1 / ln(R_o / R_i).
This checks the energy sweep, the adiabatic walls and find_Nu2 with the
mesh and metrics at three resolutions: the errors fall at second order.
The run starts from the analytical profile. DT shrinks with the mesh
(0.25, 0.12, 0.06): the NIM step damps the slowest transient at its physical
rate only while DT stays small against the finest cells' diffusion time;
twice these values take several times more steps (measured)."""
import math

import numpy as np
import pytest

import nim_amd
from nim_amd import host

pytestmark = pytest.mark.gpu

R_I, R_O = 0.3, 1.0
TOL = np.array([1e-12, 1e-12, 1e-12, 1e-6, 1e-12])


DT = {6: 0.25, 12: 0.12, 24: 0.06}


def steady_conduction(n_r):
    case = host.Case((4 * n_r, n_r, 4), R_I, R_O, 1.0, Ra=1e4, Pr=0.71, DT=DT[n_r], blend=(0.1, 0.2))
    case.g_beta = 0.0  # no buoyancy; nu keeps its Ra, Pr definition
    geo = host.build_geometry(case)
    cc = host.cell_centroids(geo)
    mid = geo["faces_mid"].reshape(-1, 3)
    T_exact = lambda x: np.log(np.hypot(x[:, 0], x[:, 1]) / R_O) / math.log(R_I / R_O)
    interior = np.ones(len(mid), dtype=bool)
    interior[geo["bc_faces"] - 1] = False
    adiabatic = geo["bc_faces"][np.isin(geo["bc_position"], (5, 6))] - 1
    st = host.initial_state(case, amp=0.0, seed=1)
    for n in ("Txyz", "Txyz_old", "Txyzt"):
        st[n] = T_exact(cc)
    ft = st["faces_T"].copy()
    ft[interior] = T_exact(mid)[interior]
    ft[adiabatic] = T_exact(mid)[adiabatic]
    st["faces_T"] = ft
    ctx = nim_amd.Context(*case.dims, case)
    ctx.upload_geometry(geo)
    ctx.build_pressure_coeffs()
    ctx.upload_state(st)
    for step in range(900):
        it, err, scal, nu = ctx.time_step(TOL, min_iter=2 if step == 0 else 0)
        if scal["dTdt_max"] < 1e-10:
            break
    assert scal["dTdt_max"] < 1e-10, scal
    for f in "uvw":
        assert np.all(ctx.get("faces_" + f) == 0.0) and np.all(ctx.get(f + "xyz") == 0.0)
    rms = host.find_rms(ctx, {"Txyz": T_exact(cc), "faces_T": T_exact(mid)}, mask={"faces_T": interior})
    ctx.close()
    return rms, nu


def test_conduction_converges_at_second_order():
    res = [steady_conduction(n) for n in (6, 12, 24)]
    nu_exact = 1.0 / math.log(R_O / R_I)
    e_cell = [r["Txyz"] for r, _ in res]
    e_face = [r["faces_T"] for r, _ in res]
    # find_Nu2 prints the inner wall as Nu_C (index 1), scaled by R_i, and
    # the outer wall as Nu_H (index 0), scaled by R_o: both are 1/ln(R_o/R_i)
    e_nu = [abs(abs(nu[1]) - nu_exact) for _, nu in res]
    e_nu_out = [abs(abs(nu[0]) - nu_exact) for _, nu in res]
    for e in (e_cell, e_face, e_nu, e_nu_out):
        orders = [math.log2(e[q] / e[q + 1]) for q in range(2)]
        assert all(o > 1.6 for o in orders), (e, orders)
    assert e_nu[-1] / nu_exact < 5e-3 and e_nu_out[-1] / nu_exact < 5e-3, (e_nu, e_nu_out)
