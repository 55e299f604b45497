// Per-cell metrics of the O-grid on the device: the heavy part of the
// reference's one-time setup (evaluate_all_mesh_constants, geometry.f90:
// 14-32), about 400 evaluations of the trilinear map per cell:
//   mapping coefficients     create_mesh.f90:254-303
//   evaluate_curvilinear     curvilinear.f90:9-160
//   volume_calculate         create_mesh.f90:306-322
//   find_normal              geometry.f90:76-190
//   calculate_diffusion_tensor_faces_i    geometry.f90:395-497
//   convection_diffusion_constants_mesh_i geometry.f90:500-632
// Inputs are the deformed nodes and outputs the reference-layout arrays, as
// the host's Fortran restatement (fortran/nim_host_geometry.f90) computes
// them; every expression keeps that order and nothing is contracted into a
// fused multiply-add, so with correctly rounded division and square root
// the results are the host's bit for bit. A face shared by two cells takes
// the higher-numbered cell's values, as the reference's serial loop leaves
// them: the cell writes its face 1, 3 (unless i = 0), 5, and its face 2, 4
// and 6 only at the last j, i and k.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "nim_hip.h"

#pragma clang fp contract(off)

int32_t nim_internal_fail(int32_t code, const char *msg);

namespace {

constexpr int NG = 5;

struct Map {
    double C[8], D[8], E[8];
};

struct Curvi {
    double x, y, z, detJ;
    double jac[3][3], jinv[3][3], g_ij[3][3], gij[3][3];
    double gam[3][3][3];  // gam[p][a][b] = Gamma^p_ab
};

__device__ void one(const double v[8], double A[8])
{
    A[0] = (v[6] + v[7] + v[4] + v[5] + v[2] + v[3] + v[0] + v[1]) / 8.;
    A[1] = (v[6] - v[7] - v[4] + v[5] + v[2] - v[3] - v[0] + v[1]) / 8.;
    A[2] = (v[6] + v[7] - v[4] - v[5] + v[2] + v[3] - v[0] - v[1]) / 8.;
    A[3] = (v[6] + v[7] + v[4] + v[5] - v[2] - v[3] - v[0] - v[1]) / 8.;
    A[4] = (v[6] - v[7] + v[4] - v[5] + v[2] - v[3] + v[0] - v[1]) / 8.;
    A[5] = (v[6] - v[7] - v[4] + v[5] - v[2] + v[3] + v[0] - v[1]) / 8.;
    A[6] = (v[6] + v[7] - v[4] - v[5] - v[2] - v[3] + v[0] + v[1]) / 8.;
    A[7] = (v[6] - v[7] + v[4] - v[5] - v[2] + v[3] - v[0] + v[1]) / 8.;
}

__device__ double det3(const double a[3][3])
{
    return a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1]
           - a[0][2] * a[1][1] * a[2][0] - a[0][1] * a[1][0] * a[2][2] - a[0][0] * a[1][2] * a[2][1];
}

// evaluate_curvilinear at (x, y, z) of the reference cube; gco(:,p) is row p
// of jac and gcon(:,p) column p of jinv
__device__ void curvilinear(const Map &m, double x, double y, double z, Curvi &cv, bool want_gamma)
{
    const double *C = m.C, *D = m.D, *E = m.E;
    cv.x = C[0] + C[1] * x + C[2] * y + C[3] * z + C[4] * x * y + C[5] * x * z + C[6] * y * z + C[7] * x * y * z;
    cv.y = D[0] + D[1] * x + D[2] * y + D[3] * z + D[4] * x * y + D[5] * x * z + D[6] * y * z + D[7] * x * y * z;
    cv.z = E[0] + E[1] * x + E[2] * y + E[3] * z + E[4] * x * y + E[5] * x * z + E[6] * y * z + E[7] * x * y * z;
    double (&a)[3][3] = cv.jac;
    a[0][0] = C[1] + C[4] * y + C[5] * z + C[7] * y * z;
    a[0][1] = D[1] + D[4] * y + D[5] * z + D[7] * y * z;
    a[0][2] = E[1] + E[4] * y + E[5] * z + E[7] * y * z;
    a[1][0] = C[2] + C[4] * x + C[6] * z + C[7] * x * z;
    a[1][1] = D[2] + D[4] * x + D[6] * z + D[7] * x * z;
    a[1][2] = E[2] + E[4] * x + E[6] * z + E[7] * x * z;
    a[2][0] = C[3] + C[5] * x + C[6] * y + C[7] * x * y;
    a[2][1] = D[3] + D[5] * x + D[6] * y + D[7] * x * y;
    a[2][2] = E[3] + E[5] * x + E[6] * y + E[7] * x * y;
    // inv3x3 (general.f90:849-857): reciprocal determinant times adjugate
    const double r = 1 / det3(a);
    double (&inv)[3][3] = cv.jinv;
    inv[0][0] = r * (a[1][1] * a[2][2] - a[2][1] * a[1][2]);
    inv[0][1] = r * (a[0][2] * a[2][1] - a[0][1] * a[2][2]);
    inv[0][2] = r * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
    inv[1][0] = r * (a[1][2] * a[2][0] - a[1][0] * a[2][2]);
    inv[1][1] = r * (a[0][0] * a[2][2] - a[0][2] * a[2][0]);
    inv[1][2] = r * (a[0][2] * a[1][0] - a[0][0] * a[1][2]);
    inv[2][0] = r * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    inv[2][1] = r * (a[0][1] * a[2][0] - a[0][0] * a[2][1]);
    inv[2][2] = r * (a[0][0] * a[1][1] - a[0][1] * a[1][0]);
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
            cv.g_ij[p][q] = a[p][0] * a[q][0] + a[p][1] * a[q][1] + a[p][2] * a[q][2];
            cv.gij[p][q] = inv[0][p] * inv[0][q] + inv[1][p] * inv[1][q] + inv[2][p] * inv[2][q];
        }
    cv.detJ = det3(a);
    if (want_gamma) {
        // find_gamma_tensor (curvilinear.f90:109-160)
        const double g12[3] = {C[4] + C[7] * z, D[4] + D[7] * z, E[4] + E[7] * z};
        const double g13[3] = {C[5] + C[7] * y, D[5] + D[7] * y, E[5] + E[7] * y};
        const double g23[3] = {C[6] + C[7] * x, D[6] + D[7] * x, E[6] + E[7] * x};
        for (int p = 0; p < 3; p++) {
            const double t12 = inv[0][p] * g12[0] + inv[1][p] * g12[1] + inv[2][p] * g12[2];
            const double t13 = inv[0][p] * g13[0] + inv[1][p] * g13[1] + inv[2][p] * g13[2];
            const double t23 = inv[0][p] * g23[0] + inv[1][p] * g23[1] + inv[2][p] * g23[2];
            cv.gam[p][0][0] = 0.; cv.gam[p][1][1] = 0.; cv.gam[p][2][2] = 0.;
            cv.gam[p][0][1] = t12; cv.gam[p][1][0] = t12;
            cv.gam[p][0][2] = t13; cv.gam[p][2][0] = t13;
            cv.gam[p][1][2] = t23; cv.gam[p][2][1] = t23;
        }
    }
}

// face_number (create_geometry.f90:626-664), 0-based
__device__ int face_number(int nt, int nr, int nz, int pos, int i, int j, int k)
{
    const int N = nt * nr * nz;
    const int o1 = N, o2 = o1 + nt * (nr - 1) * nz, o3 = o2 + nt * nr * (nz - 1), o4 = o3 + nt * nr;
    const int o5 = o4 + nt * nr, o6 = o5 + nt * nz;
    if (pos <= 2) {
        const int jj = j + (pos - 1);
        if (jj == 0) return o5 + i + nt * k;
        if (jj == nr) return o6 + i + nt * k;
        return o1 + i + nt * ((jj - 1) + (nr - 1) * k);
    }
    if (pos <= 4) {
        int ii = i + (pos - 3);
        if (ii == nt) ii = 0;
        return ii + nt * (j + nr * k);
    }
    const int kk = k + (pos - 5);
    if (kk == 0) return o3 + i + nt * j;
    if (kk == nz) return o4 + i + nt * j;
    return o2 + i + nt * (j + nr * (kk - 1));
}

struct Out {
    double *volume, *xi_x, *gij, *gij_faces, *G1, *G2, *G3;
    double *area, *normal, *tangential, *bitangential, *mid;
};

__global__ void __launch_bounds__(128) k_cell_metrics(const double *nodes, int nt, int nr, int nz, Out o,
                                                      const double *gauss)
{
    const int N = nt * nr * nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int i = c % nt, j = (c / nt) % nr, k = c / (nt * nr), ip = (i + 1) % nt;
    double gx[NG], gw[NG];
    for (int q = 0; q < NG; q++) {
        gx[q] = gauss[q];
        gw[q] = gauss[NG + q];
    }
    const int nd[8] = {i + nt * (j + (nr + 1) * k),        i + nt * (j + 1 + (nr + 1) * k),
                       ip + nt * (j + 1 + (nr + 1) * k),   ip + nt * (j + (nr + 1) * k),
                       i + nt * (j + (nr + 1) * (k + 1)),  i + nt * (j + 1 + (nr + 1) * (k + 1)),
                       ip + nt * (j + 1 + (nr + 1) * (k + 1)), ip + nt * (j + (nr + 1) * (k + 1))};
    double x[8], y[8], z[8];
    for (int a = 0; a < 8; a++) {
        x[a] = nodes[3 * nd[a] + 0];
        y[a] = nodes[3 * nd[a] + 1];
        z[a] = nodes[3 * nd[a] + 2];
    }
    Map m;
    one(x, m.C);
    one(y, m.D);
    one(z, m.E);
    Curvi cv;
    // volume_calculate
    double vol = 0.;
    for (int ii = 0; ii < NG; ii++)
        for (int jj = 0; jj < NG; jj++)
            for (int kk = 0; kk < NG; kk++) {
                curvilinear(m, gx[ii], gx[jj], gx[kk], cv, false);
                vol = vol + gw[ii] * gw[jj] * gw[kk] * (1. * cv.detJ);
            }
    o.volume[c] = vol;
    // faces this cell writes (the higher-numbered cell of each face)
    const bool writes[6] = {true, j == nr - 1, i >= 1, i == nt - 1, true, k == nz - 1};
    // find_normal: values at the six face centres
    double pts[3] = {0., 0., 0.};
    for (int f = 1; f <= 6; f++) {
        pts[0] = pts[1] = pts[2] = 0.;
        const int a = (f + 1) / 2 - 1;
        pts[a] = f % 2 == 0 ? 1. : -1.;
        if (!writes[f - 1]) continue;
        curvilinear(m, pts[0], pts[1], pts[2], cv, false);
        const int b = face_number(nt, nr, nz, f, i, j, k);
        const double sa = sqrt(cv.gij[a][a]);
        for (int q = 0; q < 3; q++) o.normal[3 * b + q] = cv.jinv[q][a] / sa;
        const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
        const double s1 = sqrt(cv.g_ij[t1][t1]), s2 = sqrt(cv.g_ij[t2][t2]);
        for (int q = 0; q < 3; q++) {
            o.tangential[3 * b + q] = cv.jac[t1][q] / s1;
            o.bitangential[3 * b + q] = cv.jac[t2][q] / s2;
        }
        o.mid[3 * b + 0] = cv.x;
        o.mid[3 * b + 1] = cv.y;
        o.mid[3 * b + 2] = cv.z;
    }
    // calculate_diffusion_tensor_faces_i: face area and face-averaged
    // g^ab / sqrt(g^aa), both accumulated over the 5 x 5 face points in the
    // reference's (ii, jj) order
    for (int f = 1; f <= 6; f++) {
        const int a = (f + 1) / 2 - 1;
        double area = 0., s[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        for (int ii = 0; ii < NG; ii++)
            for (int jj = 0; jj < NG; jj++) {
                pts[a] = f % 2 == 0 ? 1. : -1.;
                if (a == 0) { pts[1] = gx[ii]; pts[2] = gx[jj]; }
                else if (a == 1) { pts[0] = gx[ii]; pts[2] = gx[jj]; }
                else { pts[0] = gx[ii]; pts[1] = gx[jj]; }
                curvilinear(m, pts[0], pts[1], pts[2], cv, false);
                const double dS = cv.detJ * sqrt(cv.gij[a][a]);
                const double w = gw[ii] * gw[jj];
                area = area + w * dS;
                for (int bb = 0; bb < 3; bb++) {
                    const double sb = sqrt(cv.gij[bb][bb]);
                    for (int kk = 0; kk < 3; kk++) s[bb][kk] = s[bb][kk] + w * (cv.gij[bb][kk] / sb * dS);
                }
            }
        if (writes[f - 1]) o.area[face_number(nt, nr, nz, f, i, j, k)] = area;
        for (int bb = 0; bb < 3; bb++)
            for (int kk = 0; kk < 3; kk++) o.gij_faces[(size_t)c * 54 + bb + 3 * kk + 9 * (f - 1)] = s[bb][kk] / area;
    }
    // convection_diffusion_constants_mesh_i
    double acc_G[3] = {0., 0., 0.}, acc_K[3][3], acc_J[3][3];
    for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++) {
            acc_K[a][bb] = 0.;
            acc_J[a][bb] = 0.;
        }
    for (int ii = 0; ii < NG; ii++)
        for (int jj = 0; jj < NG; jj++)
            for (int kk = 0; kk < NG; kk++) {
                curvilinear(m, gx[ii], gx[jj], gx[kk], cv, true);
                const double wgt = gw[ii] * gw[jj] * gw[kk];
                for (int a = 0; a < 3; a++) {
                    const double G = cv.gam[a][0][1] * cv.gij[0][1] + cv.gam[a][1][0] * cv.gij[1][0]
                                     + cv.gam[a][0][2] * cv.gij[0][2] + cv.gam[a][2][0] * cv.gij[2][0]
                                     + cv.gam[a][1][2] * cv.gij[1][2] + cv.gam[a][2][1] * cv.gij[2][1];
                    acc_G[a] = acc_G[a] + wgt * (G * cv.detJ);
                    for (int bb = 0; bb < 3; bb++) {
                        acc_K[a][bb] = acc_K[a][bb] + wgt * (cv.gij[a][bb] * cv.detJ);
                        acc_J[a][bb] = acc_J[a][bb] + wgt * (cv.jinv[a][bb] * cv.detJ);
                    }
                }
            }
    o.G1[c] = acc_G[0] / vol;
    o.G2[c] = acc_G[1] / vol;
    o.G3[c] = acc_G[2] / vol;
    for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++) {
            o.gij[(size_t)c * 9 + a + 3 * bb] = acc_K[a][bb] / vol;
            o.xi_x[(size_t)c * 9 + a + 3 * bb] = acc_J[bb][a] / vol;
        }
}

}  // namespace

// The five-point Gauss rule in the reference's order (parameters.f90:18-22)
static void gauss_points(double x[NG], double w[NG])
{
    const double a = 1. / 3. * sqrt(5. - 2. * sqrt(10. / 7.)), b = 1. / 3. * sqrt(5. + 2. * sqrt(10. / 7.));
    x[0] = -a; x[1] = a; x[2] = 0.; x[3] = -b; x[4] = b;
    const double w1 = (322. + 13. * sqrt(70.)) / 900., w2 = (322. - 13. * sqrt(70.)) / 900.;
    w[0] = w1; w[1] = w1; w[2] = 128. / 225.; w[3] = w2; w[4] = w2;
}

extern "C" int32_t nim_device_metrics(const nim_dims *dims, const double *node_xyz, double *volume, double *xi_x,
                                      double *gij, double *gij_faces, double *G1, double *G2, double *G3,
                                      double *faces_area, double *faces_normal, double *faces_tangential,
                                      double *faces_bitangential, double *faces_mid)
{
    if (!dims || !node_xyz || !volume || !xi_x || !gij || !gij_faces || !G1 || !G2 || !G3 || !faces_area ||
        !faces_normal || !faces_tangential || !faces_bitangential || !faces_mid)
        return nim_internal_fail(NIM_E_ARG, "nim_device_metrics: null argument");
    const int nt = dims->n_theta, nr = dims->n_r, nz = dims->n_z;
    if (nt < 4 || nt % 2 || nr < 3 || nz < 3) return nim_internal_fail(NIM_E_ARG, "unsupported mesh dimensions");
    const size_t N = (size_t)nt * nr * nz, NF = 3 * N + (size_t)nt * nz + (size_t)nt * nr;
    const size_t NN = (size_t)nt * (nr + 1) * (nz + 1);
    // one device block: nodes, Gauss rule, then the outputs in argument order
    const size_t n_out[12] = {N, 9 * N, 9 * N, 54 * N, N, N, N, NF, 3 * NF, 3 * NF, 3 * NF, 3 * NF};
    size_t total = 3 * NN + 2 * NG;
    for (size_t q : n_out) total += q;
    double *dev = nullptr;
    if (hipMalloc((void **)&dev, total * sizeof(double)) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "nim_device_metrics: allocation failed");
    double g[2 * NG];
    gauss_points(g, g + NG);
    double *p = dev;
    double *d_nodes = p; p += 3 * NN;
    double *d_gauss = p; p += 2 * NG;
    double *outs[12];
    for (int q = 0; q < 12; q++) { outs[q] = p; p += n_out[q]; }
    double *host[12] = {volume, xi_x, gij, gij_faces, G1, G2, G3, faces_area, faces_normal, faces_tangential,
                        faces_bitangential, faces_mid};
    Out o{outs[0], outs[1], outs[2], outs[3], outs[4], outs[5], outs[6], outs[7], outs[8], outs[9], outs[10], outs[11]};
    int32_t rc = NIM_OK;
    if (hipMemcpy(d_nodes, node_xyz, 3 * NN * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_gauss, g, 2 * NG * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = nim_internal_fail(NIM_E_HIP, "nim_device_metrics: upload failed");
    } else {
        const int bs = 128;
        hipLaunchKernelGGL(k_cell_metrics, dim3((unsigned)((N + bs - 1) / bs)), dim3(bs), 0, 0, d_nodes, nt, nr, nz, o,
                           d_gauss);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            rc = nim_internal_fail(NIM_E_HIP, "nim_device_metrics: kernel failed");
        } else {
            for (int q = 0; q < 12 && rc == NIM_OK; q++)
                if (hipMemcpy(host[q], outs[q], n_out[q] * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                    rc = nim_internal_fail(NIM_E_HIP, "nim_device_metrics: download failed");
        }
    }
    (void)hipFree(dev);
    return rc;
}
