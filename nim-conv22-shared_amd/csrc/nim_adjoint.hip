// Discrete adjoint of the NIM time step (reverse mode, FP64, gfx950).
//
// No counterpart in the reference, which computes no sensitivities. A time
// step solves the NIM equations by the fixed-point iteration x <- G(x; p) of
// the outer passes (integrate_nse, nse.f90:117-165); p holds what the step
// takes from outside: the previous time level, the momentum coefficients of
// the step (built from the time-centred velocity), the Dirichlet wall values
// and the diffusivities nu and alpha. At the converged x* the adjoint of the
// step is the fixed point lam = G_x^T lam + xbar of the transposed pass
// (reverse accumulation of a converged iteration, Christianson 1994), and
// the step's inputs receive G_p^T lam.
//
// G_x^T lam is one pass in reverse: energy sweep, energy sources, energy
// coefficients with the adiabatic walls, pressure sweep with its wall
// updates, pressure sources, momentum sweep, momentum sources; each sweep's
// colour phases in reverse colour order and, within a cell, its updates in
// reverse order. Every stage is linearised at the converged state, which a
// pass no longer moves.
//
// Scatter: a forward unit reads neighbour values that other cells of its
// colour also read, so the reverse adds into their adjoints atomically. The
// values a unit writes are read by no other cell of its colour (the colour
// is race-free), so their adjoints are loaded, updated and stored back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "nim_hip.h"
#include "nim_stencil.h"

using namespace nim;

int32_t nim_internal_fail(int32_t code, const char *msg);
void *nim_internal_stream(nim_ctx *c);
const Dev *nim_internal_dev(nim_ctx *c);
void nim_internal_lengths(nim_ctx *c, double *R_i, double *R_o);
void nim_internal_energy_coefs(nim_ctx *c);
int32_t nim_internal_move(nim_ctx *c, double *dev, int kind, int mult, const char *name, double *host, int64_t count,
                          bool to_host);

namespace {

constexpr int BSA = 256;
// scalar parameter adjoints: partial derivatives with respect to the
// constants nu (pressure wall, momentum fits), Rinv (momentum cross
// diffusion), alpha (energy fits and cross diffusion) and the hexagon
// wall temperature
enum { P_NU = 0, P_RINV = 1, P_ALPHA = 2, P_THOT = 3, NPAR = 4 };

struct Adj {
    double *lc;    // [4][N]   cells u, v, w, T of the current iterate
    double *lf;    // [5][NF]  faces u, v, w, p, T
    double *lsrc;  // [NSRC][N] sources (consumed within a pass)
    double *lTF;   // [NCF][N]  energy coefficients (consumed within a pass)
    double *lMF;   // [NCF][N]  momentum coefficients of the step
    double *lold;  // [4][N]   previous time level
    double *ltc;   // [3][N]   time-centred velocity of the step
    double *par;   // [NPAR]   scalar parameters, this pass
    double *scr;   // [10][N]  per-cell factors of the pressure sources
};

__device__ __forceinline__ void aadd(double *p, double v) { atomicAdd(p, v); }

// sum over the block's waves into *p (every thread of the block calls it)
// (the sum over the block, one atomic per block: every thread of the block
// calls it, at most once per kernel; one address taking an atomic per wave
// of a million-thread launch serialises for hundreds of microseconds)
__device__ void par_commit(double *p, double v)
{
    __shared__ double part[BSA / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t != 0.) atomicAdd(p, t);
    }
}

// adjoints of the six faces of one unit: the faces the unit writes hold the
// adjoint of their current value and are stored back; the others collect
// this unit's contributions and are added atomically at the end
struct OwnFaces {
    int f[6];
    bool wr[6];
    double l[6];
    // (fully unrolled: the arrays stay in registers)
    __device__ void load(const double *L)
    {
#pragma unroll
        for (int p = 0; p < 6; p++) l[p] = wr[p] ? L[f[p]] : 0.;
    }
    // (the six faces are distinct: at most one matches; predicated, so that
    // l keeps static indices)
    __device__ void add(double *L, int a, double v)
    {
        bool hit = false;
#pragma unroll
        for (int p = 0; p < 6; p++) {
            const bool h = a == f[p];
            l[p] += h ? v : 0.;
            hit = hit || h;
        }
        if (!hit) aadd(L + a, v);
    }
    __device__ void store(double *L)
    {
#pragma unroll
        for (int p = 0; p < 6; p++) {
            if (wr[p]) L[f[p]] = l[p];
            else if (l[p] != 0.) aadd(L + f[p], l[p]);
        }
    }
};

// the points and weights of tder(F, fpos, i, j, k, dir): sum_m w[m] F[a[m]]
__device__ __forceinline__ void tder_pts(const Dev &d, int fpos, int i, int j, int k, int dir, int a[3], double w[3])
{
    if (dir == 0) {
        a[0] = fidx(d, fpos, ip1(d, i), j, k);
        a[1] = fidx(d, fpos, im1(d, i), j, k);
        a[2] = a[0];
        w[0] = 0.25; w[1] = -0.25; w[2] = 0.;
        return;
    }
    const int n = dir == 1 ? d.NR : d.NZ, x = dir == 1 ? j : k;
    const bool lo = x == 0, hi = x == n - 1;
    const int x0 = lo ? 0 : hi ? x : x + 1, x1 = lo ? 1 : x - 1, x2 = lo ? 2 : hi ? x - 2 : x1;
    w[0] = (lo ? -3. : hi ? 3. : 1.) / 4.;
    w[1] = (lo ? 4. : hi ? -4. : -1.) / 4.;
    w[2] = (lo ? -1. : hi ? 1. : 0.) / 4.;
    if (dir == 1) {
        a[0] = fidx(d, fpos, i, x0, k); a[1] = fidx(d, fpos, i, x1, k); a[2] = fidx(d, fpos, i, x2, k);
    } else {
        a[0] = fidx(d, fpos, i, j, x0); a[1] = fidx(d, fpos, i, j, x1); a[2] = fidx(d, fpos, i, j, x2);
    }
}

// v * tder(F, fpos, dir) into the adjoint, through the unit's own faces
// (the others atomically)
__device__ __forceinline__ void tder_add(const Dev &d, double *L, OwnFaces &o, int fpos, int i, int j, int k, int dir,
                                         double v)
{
    int a[3];
    double w[3];
    tder_pts(d, fpos, i, j, k, dir, a, w);
#pragma unroll
    for (int m = 0; m < 3; m++)
        if (w[m] != 0.) o.add(L, a[m], v * w[m]);
}

// ---------------------------------------------------------------------
// transport unit (integrate_velocity_field_formulation2_i nse.f90:831-1006,
// integrate_temperature_field_formulation2_i nse.f90:1009-1089)
// ---------------------------------------------------------------------
// Reverse of transport_unit for face field ff and cell slot cs: replaces
// the adjoints of the unit's cell and owned faces by those of their values
// before the unit, adds into the faces and cells it read and into the
// previous level, hands the coefficient adjoints of the cell to sink(slot,
// value) and returns the adjoint of the source B.
template <class Sink>
__device__ __forceinline__ double transport_vjp(const Dev &d, const Adj &a, int ff, int cs, const double *coef, Sink sink,
                                double B, double w, int i, int j, int k, int c)
{
    const int N = d.N;
    const double *Fc = d.fc + (size_t)ff * d.NF;
    const double *C = d.sc + (size_t)cs * N, *CO = d.sc + (size_t)(S_OLD + cs) * N;
    double *LF = a.lf + (size_t)ff * d.NF, *LC = a.lc + (size_t)cs * N, *LO = a.lold + (size_t)cs * N;
    const bool has2 = j < d.NR - 1, has6 = k < d.NZ - 1;
    const int jn = has2 ? j + 1 : j, kn = has6 ? k + 1 : k, ip = ip1(d, i);
    const double *gdf = d.gc + (size_t)G_DIFF * N;
    OwnFaces o;
    cell_faces(d, i, j, k, o.f);
    o.wr[0] = false; o.wr[1] = has2; o.wr[2] = false; o.wr[3] = true; o.wr[4] = false; o.wr[5] = has6;
    o.load(LF);
    // ---- forward values
    const double old = C[c], oc = CO[c];
    double av[8], bv[5], ev[5], hv[5];
    for (int m = 0; m < 8; m++) av[m] = coef[(C_F4 + m) * N + c];
    for (int m = 0; m < 5; m++) {
        bv[m] = coef[(C_F1 + m) * N + c];
        ev[m] = coef[(C_F2 + m) * N + c];
        hv[m] = coef[(C_F3 + m) * N + c];
    }
    const double F0 = Fc[o.f[0]], F1 = Fc[o.f[1]], F2 = Fc[o.f[2]], F3 = Fc[o.f[3]], F4 = Fc[o.f[4]], F5 = Fc[o.f[5]];
    const double gd0 = gdf[0 * N + c], gd1 = gdf[1 * N + c], gd2 = gdf[2 * N + c];
    const double gd3 = gdf[3 * N + c], gd4 = gdf[4 * N + c], gd5 = gdf[5 * N + c];
    const int cr = cidx(d, i, jn, k), ct = cidx(d, ip, j, k), cz = cidx(d, i, j, kn);
    const int ar2 = fidx(d, 2, i, jn, k), ar3 = fidx(d, 3, i, jn, k), ar4 = fidx(d, 4, i, jn, k);
    const int ar5 = fidx(d, 5, i, jn, k), ar6 = fidx(d, 6, i, jn, k);
    const int at1 = fidx(d, 1, ip, j, k), at2 = fidx(d, 2, ip, j, k), at4 = fidx(d, 4, ip, j, k);
    const int at5 = fidx(d, 5, ip, j, k), at6 = fidx(d, 6, ip, j, k);
    const int az1 = fidx(d, 1, i, j, kn), az2 = fidx(d, 2, i, j, kn), az3 = fidx(d, 3, i, j, kn);
    const int az4 = fidx(d, 4, i, j, kn), az6 = fidx(d, 6, i, j, kn);
    const double r2 = Fc[ar2], r3 = Fc[ar3], r4 = Fc[ar4], r5 = Fc[ar5], r6 = Fc[ar6], sr = C[cr] + CO[cr];
    const double t1 = Fc[at1], t2 = Fc[at2], t4 = Fc[at4], t5 = Fc[at5], t6 = Fc[at6], st = C[ct] + CO[ct];
    const double z1 = Fc[az1], z2 = Fc[az2], z3 = Fc[az3], z4 = Fc[az4], z6 = Fc[az6], sz = C[cz] + CO[cz];
    const double S0 = av[1] * oc + av[2] * F0 + av[3] * F1 + av[4] * F2 + av[5] * F3 + av[6] * F4 + av[7] * F5 + B;
    const double nv = (1 - w) * old + (w / av[0]) * S0, ccur = nv + oc;
    const double S2 = bv[1] * F0 + bv[2] * r2 + bv[3] * ccur + bv[4] * sr - gd0 * (F3 - F2 + r4 - r3) / 4.
                      - gd1 * (F5 - F4 + r6 - r5) / 4.;
    const double F1n = has2 ? (1 - w) * F1 + (w / bv[0]) * S2 : F1;
    const double S4 = ev[1] * F2 + ev[2] * t4 + ev[3] * ccur + ev[4] * st - gd2 * (F1n - F0 + t2 - t1) / 4.
                      - gd3 * (F5 - F4 + t6 - t5) / 4.;
    const double F3n = (1 - w) * F3 + (w / ev[0]) * S4;
    const double S6 = hv[1] * F4 + hv[2] * z6 + hv[3] * ccur + hv[4] * sz - gd4 * (F1n - F0 + z2 - z1) / 4.
                      - gd5 * (F3n - F2 + z4 - z3) / 4.;
    // ---- reverse
    double la[8] = {0, 0, 0, 0, 0, 0, 0, 0}, lb[5] = {0, 0, 0, 0, 0}, le[5] = {0, 0, 0, 0, 0}, lh[5] = {0, 0, 0, 0, 0};
    double lnv = LC[c], lccur = 0., lsr = 0., lst = 0., lsz = 0.;
    if (has6) {
        const double l = o.l[5], g = (w / hv[0]) * l;
        o.l[5] = (1 - w) * l;
        o.l[4] += g * hv[1];
        aadd(LF + az6, g * hv[2]);
        lccur += g * hv[3];
        lsz += g * hv[4];
        const double q4 = g * gd4 / 4., q5 = g * gd5 / 4.;
        o.l[1] += -q4; o.l[0] += q4; aadd(LF + az2, -q4); aadd(LF + az1, q4);
        o.l[3] += -q5; o.l[2] += q5; aadd(LF + az4, -q5); aadd(LF + az3, q5);
        lh[0] += -(g / hv[0]) * S6; lh[1] += g * F4; lh[2] += g * z6; lh[3] += g * ccur; lh[4] += g * sz;
    }
    {
        const double l = o.l[3], g = (w / ev[0]) * l;
        o.l[3] = (1 - w) * l;
        o.l[2] += g * ev[1];
        aadd(LF + at4, g * ev[2]);
        lccur += g * ev[3];
        lst += g * ev[4];
        const double q2 = g * gd2 / 4., q3 = g * gd3 / 4.;
        o.l[1] += -q2; o.l[0] += q2; aadd(LF + at2, -q2); aadd(LF + at1, q2);
        o.l[5] += -q3; o.l[4] += q3; aadd(LF + at6, -q3); aadd(LF + at5, q3);
        le[0] += -(g / ev[0]) * S4; le[1] += g * F2; le[2] += g * t4; le[3] += g * ccur; le[4] += g * st;
    }
    if (has2) {
        const double l = o.l[1], g = (w / bv[0]) * l;
        o.l[1] = (1 - w) * l;
        o.l[0] += g * bv[1];
        aadd(LF + ar2, g * bv[2]);
        lccur += g * bv[3];
        lsr += g * bv[4];
        const double q0 = g * gd0 / 4., q1 = g * gd1 / 4.;
        o.l[3] += -q0; o.l[2] += q0; aadd(LF + ar4, -q0); aadd(LF + ar3, q0);
        o.l[5] += -q1; o.l[4] += q1; aadd(LF + ar6, -q1); aadd(LF + ar5, q1);
        lb[0] += -(g / bv[0]) * S2; lb[1] += g * F0; lb[2] += g * r2; lb[3] += g * ccur; lb[4] += g * sr;
    }
    // ccur = nv + oc, partners' s = C + CO
    lnv += lccur;
    double loc = lccur;
    if (lsr != 0.) { aadd(LC + cr, lsr); aadd(LO + cr, lsr); }
    if (lst != 0.) { aadd(LC + ct, lst); aadd(LO + ct, lst); }
    if (lsz != 0.) { aadd(LC + cz, lsz); aadd(LO + cz, lsz); }
    // cell average
    const double g = (w / av[0]) * lnv;
    LC[c] = (1 - w) * lnv;
    loc += g * av[1];
    aadd(LO + c, loc);
    o.l[0] += g * av[2]; o.l[1] += g * av[3]; o.l[2] += g * av[4];
    o.l[3] += g * av[5]; o.l[4] += g * av[6]; o.l[5] += g * av[7];
    la[0] += -(g / av[0]) * S0; la[1] += g * oc; la[2] += g * F0; la[3] += g * F1; la[4] += g * F2;
    la[5] += g * F3; la[6] += g * F4; la[7] += g * F5;
    o.store(LF);
    // the FP64 sweep re-sums its diagonals from the other coefficients
    // (F4(1) = F4(2) + .. + F4(8), F(1) = F(2) + F(3) + 2 F(4) + 2 F(5)), so
    // their adjoints go to those
    for (int m = 1; m < 8; m++) la[m] += la[0];
    la[0] = 0.;
    lb[1] += lb[0]; lb[2] += lb[0]; lb[3] += 2. * lb[0]; lb[4] += 2. * lb[0]; lb[0] = 0.;
    le[1] += le[0]; le[2] += le[0]; le[3] += 2. * le[0]; le[4] += 2. * le[0]; le[0] = 0.;
    lh[1] += lh[0]; lh[2] += lh[0]; lh[3] += 2. * lh[0]; lh[4] += 2. * lh[0]; lh[0] = 0.;
    for (int m = 0; m < 8; m++) sink(C_F4 + m, la[m]);
    for (int m = 0; m < 5; m++) {
        sink(C_F1 + m, lb[m]);
        sink(C_F2 + m, le[m]);
        sink(C_F3 + m, lh[m]);
    }
    return g;
}

__device__ __forceinline__ bool phase_cell(const Dev &d, int color, int &i, int &j, int &k)
{
    return color_cell(d, color, 0, d.NZ, blockIdx.x * blockDim.x + threadIdx.x, i, j, k);
}

__device__ __forceinline__ void cell_ijk(const Dev &d, int c, int &i, int &j, int &k)
{
    i = iof(d, c % d.NT);
    j = (c / d.NT) % d.NR;
    k = c / (d.NT * d.NR);
}

// momentum sweep, one colour: one wave per velocity component
// (threadIdx.y) over 64 cells; the three components' adjoints of the
// coefficients M_F they share meet in LDS, and wave q adds the slots
// m = q mod 3 of the three into lMF
__global__ void __launch_bounds__(192) k_adj_mom_sweep(Dev d, Adj a, int color)
{
    __shared__ double part[3][NCF][64];
    int i, j, k;
    const bool act = phase_cell(d, color, i, j, k);
    const int q = threadIdx.y, t = threadIdx.x, N = d.N;
    const int c = act ? cidx(d, i, j, k) : 0;
    if (act) {
        const double w = q == 0 ? d.w_u : q == 1 ? d.w_v : d.w_w;
        const double lB = transport_vjp(d, a, q, q, d.MF, [&](int m, double v) { part[q][m][t] = v; },
                                        d.src[(R_U + q) * N + c], w, i, j, k, c);
        a.lsrc[(R_U + q) * N + c] += lB;
    }
    __syncthreads();
    if (act)
        for (int m = q; m < NCF; m += 3) a.lMF[(size_t)m * N + c] += part[0][m][t] + part[1][m][t] + part[2][m][t];
}

__global__ void __launch_bounds__(BSA) k_adj_T_sweep(Dev d, Adj a, int color)
{
    int i, j, k;
    if (!phase_cell(d, color, i, j, k)) return;
    const int N = d.N, c = cidx(d, i, j, k);
    double *LT = a.lTF;
    double lB = transport_vjp(d, a, F_T, 3, d.TF, [&](int m, double v) { LT[(size_t)m * N + c] += v; },
                              d.src[R_T * N + c], d.w_T, i, j, k, c);
    a.lsrc[R_T * N + c] += lB;
}

// ---------------------------------------------------------------------
// pressure sweep (apply_pressure_bc bc.f90:147-341 for the cell's wall
// faces, then integrate_pressure_field_i nse.f90:1091-1289)
// ---------------------------------------------------------------------
// reverse of pressure_wall for the wall face at position pos; returns the
// contribution to the adjoint of nu
__device__ __forceinline__ double pressure_wall_vjp(const Dev &d, const Adj &a, OwnFaces &o, double &lS1, int b, int pos, int i,
                                    int j, int k)
{
    const int NB = d.NB;
    const double *W = d.wall;
    double *LP = a.lf + (size_t)F_P * d.NF;
    const double n1 = W[(W_N + 0) * NB + b], n2 = W[(W_N + 1) * NB + b], n3 = W[(W_N + 2) * NB + b];
    const int fw = fidx(d, pos, i, j, k);
    const double ap = W[(W_AP + pos - 1) * NB + b], w2 = d.w_p2;
    const double l = o.l[pos - 1], g = (w2 / ap) * l;
    o.l[pos - 1] = (1 - w2) * l;
    // s = - sum over the cell's other faces of AP(q) P(q)
    for (int q = 1; q <= 6; q++)
        if (q != pos) o.add(LP, fidx(d, q, i, j, k), -g * W[(W_AP + q - 1) * NB + b]);
    // B1 = AP(8) R
    lS1 += g * W[(W_AP + 7) * NB + b];
    // - B2 - B3, tangential derivatives of p along the wall
    const double gB2 = W[W_GB2 * NB + b], gB2c = W[W_GB2C * NB + b], gB3 = W[W_GB3 * NB + b];
    if (pos <= 2) {
        tder_add(d, LP, o, pos, i, j, k, 0, -g * gB2c);
        tder_add(d, LP, o, pos, i, j, k, 2, -g * gB3);
    } else {
        tder_add(d, LP, o, pos, i, j, k, 1, -g * (j == 0 || j == d.NR - 1 ? gB2 : gB2c));
        tder_add(d, LP, o, pos, i, j, k, 0, -g * gB3);
    }
    // nu rho d2, d2 = - sum_m A(m) u_n(m) over the 5-point normal stencil
    const double *U = d.fc, *V = d.fc + d.NF, *Wv = d.fc + 2 * (size_t)d.NF;
    double *LU = a.lf, *LV = a.lf + d.NF, *LW = a.lf + 2 * (size_t)d.NF;
    double d2 = 0.;
    for (int m = 0; m < 5; m++) {
        const int q = d.wpts[m * NB + b];
        const double am = W[(W_A + m) * NB + b];
        d2 -= am * (U[q] * n1 + V[q] * n2 + Wv[q] * n3);
        const double s = -g * d.nu * d.rho * am;
        if (s != 0.) {
            aadd(LU + q, s * n1);
            aadd(LV + q, s * n2);
            aadd(LW + q, s * n3);
        }
    }
    // - b_n rho, b_n = - g_beta T_w n3
    if (n3 != 0.) aadd(a.lf + (size_t)F_T * d.NF + fw, g * d.g_beta * d.rho * n3);
    return g * d.rho * d2;
}

__global__ void __launch_bounds__(BSA) k_adj_ppe_sweep(Dev d, Adj a, int color)
{
    int i, j, k;
    double pnu = 0.;
    if (phase_cell(d, color, i, j, k)) {
        const int N = d.N, NT = d.NT, NR = d.NR, NZ = d.NZ;
        const int c = cidx(d, i, j, k);
        double *LP = a.lf + (size_t)F_P * d.NF, *LS = a.lsrc + (size_t)R_P * N;
        const double *gdf = d.gc + (size_t)G_DIFF * N;
        const double w = d.w_p;
        const bool hx = j < NR - 1, hz = k < NZ - 1;
        const int jn = hx ? j + 1 : j, kn = hz ? k + 1 : k, ip = ip1(d, i);
        OwnFaces o;
        cell_faces(d, i, j, k, o.f);
        o.wr[0] = j == 0; o.wr[1] = true; o.wr[2] = false; o.wr[3] = true; o.wr[4] = k == 0; o.wr[5] = true;
        o.load(LP);
        const double *Q = d.AQ;
        double qx[NAQ], qy[NAQ], qz[NAQ];
        for (int m = 0; m < NAQ; m++) {
            qx[m] = Q[m * N + c];
            qy[m] = Q[(NAQ + m) * N + c];
            qz[m] = Q[(2 * NAQ + m) * N + c];
        }
        double lS1 = 0.;
        // z face (face 6)
        if (hz) {
            const double l = o.l[5], g = (w / qz[Q_DIAG]) * l;
            o.l[5] = (1 - w) * l;
            o.add(LP, o.f[0], -g * qz[Q_TA0]); o.add(LP, o.f[1], -g * qz[Q_TA0]);
            o.add(LP, fidx(d, 1, i, j, kn), -g * qz[Q_TB0]); o.add(LP, fidx(d, 2, i, j, kn), -g * qz[Q_TB0]);
            o.add(LP, o.f[2], -g * qz[Q_TA1]); o.add(LP, o.f[3], -g * qz[Q_TA1]);
            o.add(LP, fidx(d, 3, i, j, kn), -g * qz[Q_TB1]); o.add(LP, fidx(d, 4, i, j, kn), -g * qz[Q_TB1]);
            o.add(LP, o.f[4], -g * qz[Q_SELF]); o.add(LP, fidx(d, 6, i, j, kn), -g * qz[Q_FAR]);
            lS1 += g * qz[Q_S1];
            aadd(LS + cidx(d, i, j, kn), g * qz[Q_S2]);
            tder_add(d, LP, o, 6, i, j, k, 1, g * gdf[4 * N + c]);
            tder_add(d, LP, o, 6, i, j, k, 0, g * gdf[5 * N + c]);
        }
        // y face (face 4)
        {
            const double l = o.l[3], g = (w / qy[Q_DIAG]) * l;
            o.l[3] = (1 - w) * l;
            o.add(LP, o.f[0], -g * qy[Q_TA0]); o.add(LP, o.f[1], -g * qy[Q_TA0]);
            o.add(LP, fidx(d, 1, ip, j, k), -g * qy[Q_TB0]); o.add(LP, fidx(d, 2, ip, j, k), -g * qy[Q_TB0]);
            o.add(LP, o.f[2], -g * qy[Q_SELF]); o.add(LP, fidx(d, 4, ip, j, k), -g * qy[Q_FAR]);
            o.add(LP, o.f[4], -g * qy[Q_TA1]); o.add(LP, o.f[5], -g * qy[Q_TA1]);
            o.add(LP, fidx(d, 5, ip, j, k), -g * qy[Q_TB1]); o.add(LP, fidx(d, 6, ip, j, k), -g * qy[Q_TB1]);
            lS1 += g * qy[Q_S1];
            aadd(LS + cidx(d, ip, j, k), g * qy[Q_S2]);
            tder_add(d, LP, o, 4, i, j, k, 1, g * gdf[2 * N + c]);
            tder_add(d, LP, o, 4, i, j, k, 2, g * gdf[3 * N + c]);
        }
        // x face (face 2)
        if (hx) {
            const double l = o.l[1], g = (w / qx[Q_DIAG]) * l;
            o.l[1] = (1 - w) * l;
            o.add(LP, o.f[0], -g * qx[Q_SELF]); o.add(LP, fidx(d, 2, i, jn, k), -g * qx[Q_FAR]);
            o.add(LP, o.f[2], -g * qx[Q_TA0]); o.add(LP, o.f[3], -g * qx[Q_TA0]);
            o.add(LP, fidx(d, 3, i, jn, k), -g * qx[Q_TB0]); o.add(LP, fidx(d, 4, i, jn, k), -g * qx[Q_TB0]);
            o.add(LP, o.f[4], -g * qx[Q_TA1]); o.add(LP, o.f[5], -g * qx[Q_TA1]);
            o.add(LP, fidx(d, 5, i, jn, k), -g * qx[Q_TB1]); o.add(LP, fidx(d, 6, i, jn, k), -g * qx[Q_TB1]);
            lS1 += g * qx[Q_S1];
            aadd(LS + cidx(d, i, jn, k), g * qx[Q_S2]);
            tder_add(d, LP, o, 2, i, j, k, 0, g * gdf[0 * N + c]);
            tder_add(d, LP, o, 2, i, j, k, 2, g * gdf[1 * N + c]);
        }
        // walls, reverse of bottom, top, hexagon, cylinder
        if (j == NR - 1) pnu += pressure_wall_vjp(d, a, o, lS1, 2 * NT * NR + NT * NZ + i + NT * k, 2, i, j, k);
        if (j == 0) pnu += pressure_wall_vjp(d, a, o, lS1, 2 * NT * NR + i + NT * k, 1, i, j, k);
        if (k == NZ - 1) pnu += pressure_wall_vjp(d, a, o, lS1, NT * NR + i + NT * j, 6, i, j, k);
        if (k == 0) pnu += pressure_wall_vjp(d, a, o, lS1, i + NT * j, 5, i, j, k);
        o.store(LP);
        aadd(LS + c, lS1);
    }
    par_commit(a.par + P_NU, pnu);
}

// ---------------------------------------------------------------------
// sources: the transposed stencils gathered per face (no atomics)
// ---------------------------------------------------------------------
// The cells along a non-periodic direction of n cells whose tder stencil
// (three_point: central (1, -1, 0) / 4 on (x + 1, x - 1), low wall
// (-3, 4, -1) / 4 on (0, 1, 2), high wall (3, -4, 1) / 4 on (n - 1, n - 2,
// n - 3)) reads the value at index x, and the weights they read it with: four
// candidates x - 1, x + 1 (central readers), 0 and n - 1 (wall readers),
// weight 0 (at a valid index) where a candidate does not read x
__device__ __forceinline__ void row_readers(int x, int n, int r[4], double w[4])
{
    r[0] = x - 1 < 0 ? 0 : x - 1;
    w[0] = (x - 1 >= 1 && x - 1 <= n - 2) ? 0.25 : 0.;
    r[1] = x + 1 > n - 1 ? n - 1 : x + 1;
    w[1] = (x + 1 >= 1 && x + 1 <= n - 2) ? -0.25 : 0.;
    r[2] = 0;
    w[2] = x <= 2 ? (x == 0 ? -3. : x == 1 ? 4. : -1.) / 4. : 0.;
    r[3] = n - 1;
    w[3] = x >= n - 3 ? (x == n - 1 ? 3. : x == n - 2 ? -4. : 1.) / 4. : 0.;
}

// Face slot s of a face field: its family (0 radial, 1 azimuthal, 2 axial)
// and indices. Radial R(i, x, k) lies between cells (i, x - 1, k) (its
// face 2) and (i, x, k) (its face 1); azimuthal T(x, j, k) between
// (x - 1, j, k) (face 4) and (x, j, k) (face 3); axial Z(i, j, x) between
// (i, j, x - 1) (face 6) and (i, j, x) (face 5).
struct FaceAt {
    int fam, i, j, k;  // the coordinate across the face is the family's own (j, i or k)
};

__device__ __forceinline__ FaceAt face_at(const Dev &d, int s)
{
    FaceAt f;
    const int NT = d.NT, NR = d.NR;
    if (s < d.oT) {
        f.fam = 0; f.i = iof(d, s % NT); f.j = (s / NT) % (NR + 1); f.k = s / (NT * (NR + 1));
    } else if (s < d.oZ) {
        const int t = s - d.oT;
        f.fam = 1; f.i = iof(d, t % NT); f.j = (t / NT) % NR; f.k = t / (NT * NR);
    } else {
        const int t = s - d.oZ;
        f.fam = 2; f.i = iof(d, t % NT); f.j = (t / NT) % NR; f.k = t / (NT * NR);
    }
    return f;
}

// The two cells of a face (-1 where a wall leaves one) and the face
// position in each: lo is the cell below (the face is its 2, 4 or 6), hi
// the cell above (its 1, 3 or 5)
__device__ __forceinline__ void face_cells(const Dev &d, const FaceAt &f, int &lo, int &hi)
{
    if (f.fam == 0) {
        lo = f.j > 0 ? cidx(d, f.i, f.j - 1, f.k) : -1;
        hi = f.j < d.NR ? cidx(d, f.i, f.j, f.k) : -1;
    } else if (f.fam == 1) {
        lo = cidx(d, im1(d, f.i), f.j, f.k);
        hi = cidx(d, f.i, f.j, f.k);
    } else {
        lo = f.k > 0 ? cidx(d, f.i, f.j, f.k - 1) : -1;
        hi = f.k < d.NZ ? cidx(d, f.i, f.j, f.k) : -1;
    }
}

// The transpose of the cross-diffusion operator X (cross_diffusion with
// coef = 1) at face f, as the cells whose X reads the face and the weight
// each reads it with: X^T lam at the face is the sum of lam(cell) times
// weight. A fixed 18 entries, unused ones with weight 0 at a valid cell, so
// that every load issues unconditionally:
//  * the normal derivative G(own family) (F(face 2/4/6) - F(face 1/3/5)) / 2
//    of the cell below the face (+G/2) and above it (-G/2);
//  * the cross terms -GO(m) (tder(face 2/4/6, dir) - tder(face 1/3/5, dir)) / 2
//    of the two directions along the face, in the rows of both cells: the
//    cells whose tder stencil reads the face (theta: i -+ 1 at +-1/4, two of
//    the four entries; r and z: row_readers), times -GO/2 in the row below
//    and +GO/2 in the row above.
struct Readers {
    int c[18];
    double w[18];
};

__device__ __forceinline__ void face_readers(const Dev &d, const FaceAt &f, Readers &R)
{
    const int N = d.N, NR = d.NR, NZ = d.NZ;
    const double *g = d.gc;
    // own-family G slot; the two directions along the face, their GO slots
    // (fxy, fxz / fyx, fyz / fzy, fzx) and directions (0 theta, 1 r, 2 z)
    int Gslot, m1, dir1, m2, dir2;
    if (f.fam == 0) { Gslot = G_G + 0; m1 = 0; dir1 = 0; m2 = 2; dir2 = 2; }
    else if (f.fam == 1) { Gslot = G_G + 1; m1 = 1; dir1 = 1; m2 = 4; dir2 = 2; }
    else { Gslot = G_G + 2; m1 = 5; dir1 = 0; m2 = 3; dir2 = 1; }
    // the cell below (side 0) and above (side 1), clamped into the grid where
    // a wall leaves one out
    bool ok[2] = {true, true};
    int ci[2] = {f.i, f.i}, cj[2] = {f.j, f.j}, ck[2] = {f.k, f.k};
    if (f.fam == 0) {
        ok[0] = f.j > 0; ok[1] = f.j < NR;
        cj[0] = f.j > 0 ? f.j - 1 : 0; cj[1] = f.j < NR ? f.j : NR - 1;
    } else if (f.fam == 1) {
        ci[0] = im1(d, f.i);
    } else {
        ok[0] = f.k > 0; ok[1] = f.k < NZ;
        ck[0] = f.k > 0 ? f.k - 1 : 0; ck[1] = f.k < NZ ? f.k : NZ - 1;
    }
    int n = 0;
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const int c = cidx(d, ci[side], cj[side], ck[side]);
        const double G = g[(size_t)Gslot * N + c];
        R.c[n] = c;
        R.w[n++] = ok[side] ? G * (side == 0 ? 0.5 : -0.5) : 0.;
    }
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const double sgn = side == 0 ? -0.5 : 0.5;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int m = t == 0 ? m1 : m2, dir = t == 0 ? dir1 : dir2;
            const double *GO = g + (size_t)(G_GO + m) * N;
            int r[4], cc[4];
            double w[4];
            if (dir == 0) {
                cc[0] = cidx(d, im1(d, ci[side]), cj[side], ck[side]);
                cc[1] = cidx(d, ip1(d, ci[side]), cj[side], ck[side]);
                cc[2] = cc[0];
                cc[3] = cc[1];
                w[0] = 0.25; w[1] = -0.25; w[2] = 0.; w[3] = 0.;
            } else {
                if (dir == 1) row_readers(cj[side], NR, r, w);
                else row_readers(ck[side], NZ, r, w);
#pragma unroll
                for (int q = 0; q < 4; q++)
                    cc[q] = dir == 1 ? cidx(d, ci[side], r[q], ck[side]) : cidx(d, ci[side], cj[side], r[q]);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double go = GO[cc[q]];
                R.c[n] = cc[q];
                R.w[n++] = ok[side] ? go * (sgn * w[q]) : 0.;
            }
        }
    }
}

__device__ __forceinline__ double readers_sum(const Readers &R, const double *lam)
{
    double S = 0.;
#pragma unroll
    for (int n = 0; n < 18; n++) S += lam[R.c[n]] * R.w[n];
    return S;
}

// d out / d phi(face) of dphi(phi, comp) = invV (sum of +-phi A_comp) of
// each cell of the face, times fac(cell, comp): the face is the cell's
// +side face (2, 4, 6) in the lo cell and its -side face in the hi cell.
// A missing cell (wall) is replaced by the other one with weight 0.
template <class Fac>
__device__ __forceinline__ double dphi_T(const Dev &d, const FaceAt &f, int s, Fac fac)
{
    int lo, hi;
    face_cells(d, f, lo, hi);
    const bool lok = lo >= 0, hok = hi >= 0;
    const int l = lok ? lo : hi, h = hok ? hi : lo;
    const double vl = d.gc[G_INVV * d.N + l], vh = d.gc[G_INVV * d.N + h];
    double r = 0.;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const double A = d.An[(size_t)comp * d.NF + s];
        const double fl = fac(l, comp), fh = fac(h, comp);
        r += (lok ? fl * vl * A : 0.) - (hok ? fh * vh * A : 0.);
    }
    return r;
}

// update_velocity_sources (mom_src): R = Rinv X(F_q) + dphi_q(p) / rho + b
// (b = -g_beta (T + T_old) / 2 for w); one thread per face slot
// (grid-stride over the faces, so that the parameter sum takes one atomic
// per block of a bounded grid)
__global__ void __launch_bounds__(BSA) k_adj_mom_src_faces(Dev d, Adj a)
{
    double prinv = 0.;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < d.NF; s += gridDim.x * blockDim.x) {
        const int N = d.N, NF = d.NF;
        const FaceAt f = face_at(d, s);
        Readers R;
        face_readers(d, f, R);
        double S[3];
#pragma unroll
        for (int q = 0; q < 3; q++) S[q] = readers_sum(R, a.lsrc + (size_t)(R_U + q) * N);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            a.lf[(size_t)q * NF + s] += d.Rinv * S[q];
            prinv += d.fc[(size_t)q * NF + s] * S[q];
        }
        const double *L = a.lsrc;
        a.lf[(size_t)F_P * NF + s] +=
            dphi_T(d, f, s, [&](int c, int comp) { return L[(size_t)(R_U + comp) * N + c] * d.rinv; });
    }
    par_commit(a.par + P_RINV, prinv);
}

__global__ void __launch_bounds__(BSA) k_adj_mom_src_cells(Dev d, Adj a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.N) return;
    const int N = d.N;
    const double s = -(d.g_beta * 0.5) * a.lsrc[(R_U + 2) * N + c];
    a.lc[3 * N + c] += s;
    a.lold[3 * N + c] += s;
    for (int q = 0; q < 3; q++) a.lsrc[(R_U + q) * N + c] = 0.;
}

// update_pressure_sources (ppe_src): R_D_p + R_b_p. The cells' factors
// first: R_b_p = -(ux^2 + vy^2 + wz^2) - 2 (uy vx + vz wy + wx uz) - bzz
// + (ux + vy + wz) / (2 DT), bzz = dphi_z(-g_beta T), so with l the source
// adjoint: d/d(u_x, u_y, u_z) = l (-2 ux + 1/(2DT), -2 vx, -2 wx), and
// likewise for v and w; the T faces get l g_beta through dphi_z
__global__ void __launch_bounds__(BSA) k_adj_ppe_factors(Dev d, Adj a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.N) return;
    const int N = d.N;
    const double l = a.lsrc[R_P * N + c];
    int i, j, k;
    cell_ijk(d, c, i, j, k);
    int f[6];
    cell_faces(d, i, j, k, f);
    const double *U = d.fc, *V = d.fc + d.NF, *W = d.fc + 2 * (size_t)d.NF;
    double uf[6], vf[6], wf[6];
    for (int p = 0; p < 6; p++) {
        uf[p] = U[f[p]];
        vf[p] = V[f[p]];
        wf[p] = W[f[p]];
    }
    const double invV = d.gc[G_INVV * N + c];
    const double ux = dphi(d, d.An, uf, f, 0, invV), uy = dphi(d, d.An, uf, f, 1, invV), uz = dphi(d, d.An, uf, f, 2, invV);
    const double vx = dphi(d, d.An, vf, f, 0, invV), vy = dphi(d, d.An, vf, f, 1, invV), vz = dphi(d, d.An, vf, f, 2, invV);
    const double wx = dphi(d, d.An, wf, f, 0, invV), wy = dphi(d, d.An, wf, f, 1, invV), wz = dphi(d, d.An, wf, f, 2, invV);
    const double hD = 1. / (2. * d.DT);
    double *F = a.scr;
    F[0 * N + c] = l * (-2. * ux + hD); F[1 * N + c] = l * (-2. * vx); F[2 * N + c] = l * (-2. * wx);
    F[3 * N + c] = l * (-2. * uy); F[4 * N + c] = l * (-2. * vy + hD); F[5 * N + c] = l * (-2. * wy);
    F[6 * N + c] = l * (-2. * uz); F[7 * N + c] = l * (-2. * vz); F[8 * N + c] = l * (-2. * wz + hD);
    F[9 * N + c] = l * d.g_beta;
}

__global__ void __launch_bounds__(BSA) k_adj_ppe_src_faces(Dev d, Adj a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.NF) return;
    const int N = d.N, NF = d.NF;
    const FaceAt f = face_at(d, s);
    Readers R;
    face_readers(d, f, R);
    a.lf[(size_t)F_P * NF + s] += d.rinv * readers_sum(R, a.lsrc + (size_t)R_P * N);
    const double *F = a.scr;
    for (int q = 0; q < 3; q++)
        a.lf[(size_t)q * NF + s] += dphi_T(d, f, s, [&](int c, int comp) { return F[(size_t)(3 * q + comp) * N + c]; });
    a.lf[(size_t)F_T * NF + s] += dphi_T(d, f, s, [&](int c, int comp) { return comp == 2 ? F[(size_t)9 * N + c] : 0.; });
}

// update_temperature_sources (T_src): alpha X(T)
__global__ void __launch_bounds__(BSA) k_adj_T_src_faces(Dev d, Adj a)
{
    double palpha = 0.;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < d.NF; s += gridDim.x * blockDim.x) {
        const FaceAt f = face_at(d, s);
        Readers R;
        face_readers(d, f, R);
        const double S = readers_sum(R, a.lsrc + (size_t)R_T * d.N);
        a.lf[(size_t)F_T * d.NF + s] += d.alpha * S;
        palpha += d.fc[(size_t)F_T * d.NF + s] * S;
    }
    par_commit(a.par + P_ALPHA, palpha);
}

// the consumed source adjoints
__global__ void __launch_bounds__(BSA) k_adj_clear_src(Dev d, Adj a, int slot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < d.N) a.lsrc[(size_t)slot * d.N + c] = 0.;
}

// ---------------------------------------------------------------------
// coefficients (nim.f90:160-330 momentum, 334-503 energy) and the adiabatic
// walls (apply_zero_flux, bc.f90:345-390)
// ---------------------------------------------------------------------
// value and partial derivatives with respect to the contravariant velocity
// and to k of the six exponential-fit constants of one direction
__device__ void direction_d(const Dev &d, int c, int dir, double ux, double uy, double uz, double kd, double A[6],
                            double dAu[6], double dAk[6], double &vel, double &kk)
{
    const int N = d.N;
    const double *g = d.gc;
    vel = ux * g[(G_XI + 3 * dir + 0) * N + c] + uy * g[(G_XI + 3 * dir + 1) * N + c]
          + uz * g[(G_XI + 3 * dir + 2) * N + c];
    kk = g[(G_GD + dir) * N + c] * kd;
    const double gp = g[(G_GF + 2 * dir + 1) * N + c], gm = g[(G_GF + 2 * dir) * N + c];
    Dual R[6];
    nim_direction_t<Dual>(Dual{vel, 1.}, Dual{kk, 0.}, gp, gm, d.v_limit, R);
    for (int m = 0; m < 6; m++) {
        A[m] = R[m].v;
        dAu[m] = R[m].d;
    }
    nim_direction_t<Dual>(Dual{vel, 0.}, Dual{kk, 1.}, gp, gm, d.v_limit, R);
    for (int m = 0; m < 6; m++) dAk[m] = R[m].d;
}

// the velocity of the fits at cell c: time-centred (momentum) or the mean of
// the current iterate and the previous level (energy)
__device__ void fit_velocity(const Dev &d, int c, int energy, double u[3])
{
    const int N = d.N;
    for (int q = 0; q < 3; q++)
        u[q] = energy ? 0.5 * (d.sc[(S_CUR + q) * N + c] + d.sc[(S_OLD + q) * N + c]) : d.sc[(S_TC + q) * N + c];
}

// reverse of k_coef_A + k_coef_F (energy = 0, adjoints from lMF into the
// time-centred velocity) or of T_coef (energy = 1, from lTF into the
// current and previous velocity, then the adiabatic walls); the diffusivity
// part goes to par[pk]. Gathered per cell: the constants A of a cell's
// direction d enter its own F4 and F(d) and, as B, the F(d) of its -
// neighbour in d (nim.f90:304-327). The cell's thread takes the adjoints of
// all three, evaluates A and its derivatives once per direction and writes
// only its own cell. The coefficient adjoints it consumes are cleared by
// k_adj_coef_clear afterwards: the + neighbour reads them here too.
__global__ void __launch_bounds__(BSA) k_adj_coef(Dev d, Adj a, int energy, double *par, int pk)
{
    double pkd = 0.;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.N; c += gridDim.x * blockDim.x) {
        const int N = d.N;
        int i, j, k;
        cell_ijk(d, c, i, j, k);
        const double *LF = energy ? a.lTF : a.lMF;
        const double kd = energy ? d.alpha : d.nu;
        // the forward wrote F(d) of this cell where its + neighbour exists,
        // and F(d) of the - neighbour where that exists
        const bool hasp[3] = {j < d.NR - 1, true, k < d.NZ - 1};
        const bool hasm[3] = {j > 0, true, k > 0};
        const int cm[3] = {c - d.NT, cidx(d, im1(d, i), j, k), c - d.NT * d.NR};
        const int base[3] = {C_F1, C_F2, C_F3};
        double l4[8];
        for (int m = 0; m < 8; m++) l4[m] = LF[(C_F4 + m) * N + c];
        const double lf2 = l4[0] + l4[1];
        double lu[3] = {0., 0., 0.}, u[3];
        bool fitted = false;
        for (int dir = 0; dir < 3; dir++) {
            const int s = base[dir];
            const double lfa = l4[0] + l4[2 + 2 * dir], lfb = l4[0] + l4[3 + 2 * dir];
            double lb = 0., lh = 0., le = 0., lq = 0.;
            if (hasp[dir]) {
                const double f0 = LF[s * N + c];
                lb = f0 + LF[(s + 1) * N + c];
                lh = 2. * f0 + LF[(s + 3) * N + c];
            }
            if (hasm[dir]) {
                const int m = cm[dir];
                const double f0 = LF[s * N + m];
                le = f0 + LF[(s + 2) * N + m];
                lq = 2. * f0 + LF[(s + 4) * N + m];
            }
            if (lf2 == 0. && lfa == 0. && lfb == 0. && lb == 0. && lh == 0. && le == 0. && lq == 0.) continue;
            if (!fitted) {
                fit_velocity(d, c, energy, u);
                fitted = true;
            }
            double A[6], dAu[6], dAk[6], vel, kk;
            direction_d(d, c, dir, u[0], u[1], u[2], kd, A, dAu, dAk, vel, kk);
            const double A0 = A[0], A1 = A[1], A4 = A[4], A5 = A[5], A44 = A4 * A4;
            double lA[6] = {0., 0., 0., 0., 0., 0.};
            // F4: f2 = (1/A14 + 1/A24 + 1/A34 - 1/DT) / 2, f(3+2d) = -A(d)5 / A(d)4,
            // f(4+2d) = (A(d)5 - 1) / A(d)4, F4 = [f2 + ... + f8, f2, ..., f8]
            lA[4] += lf2 * 0.5 * (-1. / A44) + lfa * (A5 / A44) - lfb * ((A5 - 1.) / A44);
            lA[5] += -lfa / A4 + lfb / A4;
            // F(d) = [b + e + 2 h + 2 q, b, e, h, q] with b = A2 - A0 A5 / A4,
            // h = A0 / (2 A4) of this cell, e = B3 - B1 / B4 + B1 B5 / B4,
            // q = B1 / (2 B4) of its + neighbour: here b, h of this cell's
            // F(d) and e, q of the - neighbour's
            lA[2] += lb;
            lA[0] += -lb * A5 / A4 + lh / (2. * A4);
            lA[5] += -lb * A0 / A4;
            lA[4] += lb * A0 * A5 / A44 - lh * A0 / (2. * A44);
            lA[3] += le;
            lA[1] += le * (-1. / A4 + A5 / A4) + lq / (2. * A4);
            lA[4] += le * (A1 / A44 - A1 * A5 / A44) - lq * A1 / (2. * A44);
            lA[5] += le * A1 / A4;
            double lvel = 0., lk = 0.;
            for (int m = 0; m < 6; m++) {
                lvel += lA[m] * dAu[m];
                lk += lA[m] * dAk[m];
            }
            for (int q = 0; q < 3; q++) lu[q] += lvel * d.gc[(G_XI + 3 * dir + q) * N + c];
            pkd += lk * d.gc[(G_GD + dir) * N + c];
        }
        for (int q = 0; q < 3; q++) {
            if (lu[q] == 0.) continue;
            if (energy) {
                a.lc[q * N + c] += 0.5 * lu[q];
                a.lold[q * N + c] += 0.5 * lu[q];
            } else {
                a.ltc[q * N + c] += lu[q];
            }
        }
        // adiabatic top and bottom walls: T_w = (Z0 T_a - Z1 T_b + Z2 T_c) / Z3
        if (energy && (k == 0 || k == d.NZ - 1)) {
            double *LT = a.lf + (size_t)F_T * d.NF;
            const double *W = d.wall;
            const int NB = d.NB;
            int b, fw, fa, fb, fcc;
            if (k == 0) {
                b = i + d.NT * j;
                fw = fidx(d, 5, i, j, 0); fa = fidx(d, 6, i, j, 0); fb = fidx(d, 6, i, j, 1); fcc = fidx(d, 6, i, j, 2);
            } else {
                b = d.NT * d.NR + i + d.NT * j;
                fw = fidx(d, 6, i, j, k); fa = fidx(d, 5, i, j, k); fb = fidx(d, 5, i, j, k - 1); fcc = fidx(d, 5, i, j, k - 2);
            }
            const double l = LT[fw] / W[(W_ZF + 3) * NB + b];
            LT[fw] = 0.;
            aadd(LT + fa, l * W[(W_ZF + 0) * NB + b]);
            aadd(LT + fb, -l * W[(W_ZF + 1) * NB + b]);
            aadd(LT + fcc, l * W[(W_ZF + 2) * NB + b]);
        }
    }
    par_commit(par + pk, pkd);
}

// the coefficient adjoints k_adj_coef consumed: F4 and, where the forward
// wrote them (the + neighbour exists), F1..F3; the others keep theirs
__global__ void __launch_bounds__(BSA) k_adj_coef_clear(Dev d, Adj a, int energy)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.N) return;
    const int N = d.N, j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    double *LF = energy ? a.lTF : a.lMF;
    const bool nb[3] = {j < d.NR - 1, true, k < d.NZ - 1};
    for (int m = 0; m < NCF; m++)
        if (m < C_F1 || nb[(m - C_F1) / 5]) LF[(size_t)m * N + c] = 0.;
}

// ---------------------------------------------------------------------
// pass and step bookkeeping
// ---------------------------------------------------------------------
// Dirichlet slot of a face field: u, v, w on every wall, T on the hexagon
// and cylinder (the adiabatic top and bottom are unknowns), p on none
__device__ __forceinline__ bool dirichlet(const Dev &d, int field, int s)
{
    if (field == F_P) return false;
    const int NT = d.NT, NR = d.NR, NZ = d.NZ;
    if (s < d.oT) {
        const int j = (s / NT) % (NR + 1);
        return j == 0 || j == NR;
    }
    if (s < d.oZ) return false;
    if (field == F_T) return false;
    const int k = (s - d.oZ) / (NT * NR);
    return k == 0 || k == NZ;
}

// start of a reverse pass: clear the intermediates and the parameter
// adjoints of the pass, and the adjoints of the Dirichlet face values
__global__ void __launch_bounds__(BSA) k_adj_begin(Dev d, Adj a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    const long N = d.N, NF = d.NF;
    for (long s = t; s < NSRC * N; s += stride) a.lsrc[s] = 0.;
    for (long s = t; s < NCF * N; s += stride) { a.lTF[s] = 0.; a.lMF[s] = 0.; }
    for (long s = t; s < 4 * N; s += stride) a.lold[s] = 0.;
    for (long s = t; s < 5 * NF; s += stride)
        if (dirichlet(d, (int)(s / NF), (int)(s % NF))) a.lf[s] = 0.;
    if (t < NPAR) a.par[t] = 0.;
}

// lam += seed over cells and faces; delta and magnitude maxima of the
// change against prev (bit patterns of non-negative doubles, atomic max),
// then prev = lam
__global__ void __launch_bounds__(BSA) k_adj_add_seed(double *lam, const double *seed, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) lam[s] += seed[s];
}

__global__ void __launch_bounds__(BSA) k_adj_delta(double *lam, double *prev, long n, unsigned long long *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    double dm = 0., lm = 0.;
    for (long s = t; s < n; s += stride) {
        const double x = lam[s];
        dm = fmax(dm, fabs(x - prev[s]));
        lm = fmax(lm, fabs(x));
        prev[s] = x;
    }
    dm = wave_max(dm);
    lm = wave_max(lm);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, (unsigned long long)__double_as_longlong(dm));
        atomicMax(out + 1, (unsigned long long)__double_as_longlong(lm));
    }
}

// the seed of a reverse step from the adjoint of the state after it:
// cells cbar plus half the time-centred adjoint (TC' = (x + x') / 2), faces
// fbar; lam = prev = seed
__global__ void __launch_bounds__(BSA) k_adj_seed(Dev d, double *seed, double *lam, double *prev, const double *cbar,
                                                  const double *tcbar, const double *fbar)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    const long N = d.N, n = 4 * N + 5 * (long)d.NF;
    for (long s = t; s < n; s += stride) {
        double v;
        if (s < 4 * N) v = cbar[s] + (s < 3 * N ? 0.5 * tcbar[s] : 0.);
        else v = fbar[s - 4 * N];
        seed[s] = v;
        lam[s] = v;
        prev[s] = v;
    }
}

// after a converged reverse step: the hexagon wall temperatures and the
// scalar parameters of the last pass into theta, then the adjoint of the
// state before the step: cells = previous level + half the time-centred
// adjoint of the state after the step, time-centred velocity from the
// momentum coefficients (ltc), faces 0 (they only start the next solve)
__global__ void __launch_bounds__(BSA) k_adj_hex(Dev d, Adj a, double *theta, double *thf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.;
    if (t < d.NT * d.NZ) {
        const int i = t % d.NT, k = t / d.NT;
        s = a.lf[(size_t)F_T * d.NF + fidx(d, 1, i, 0, k)];
        thf[t] += s;  // hexagon face t = i + n_theta k, the reference's boundary order
    }
    par_commit(theta + P_THOT, s);
}

__global__ void __launch_bounds__(BSA) k_adj_carry(Dev d, Adj a, double *cbar, double *tcbar, double *fbar)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    const long N = d.N, NF = d.NF;
    for (long s = t; s < 4 * N; s += stride) {
        cbar[s] = a.lold[s] + (s < 3 * N ? 0.5 * tcbar[s] : 0.);
    }
    for (long s = t; s < 3 * N; s += stride) tcbar[s] = a.ltc[s];
    for (long s = t; s < 5 * NF; s += stride) fbar[s] = 0.;
}

__global__ void k_adj_par_add(double *theta, const double *par)
{
    if (threadIdx.x < P_THOT) theta[threadIdx.x] += par[threadIdx.x];
}

// adjoint of find_Nu2 (nse.f90:2699-2798; k_nusselt): J = wc Nu_cyl + wh Nu_hex
// into the face adjoint fbar, and its direct part through the hexagon wall
// temperature into theta
__global__ void __launch_bounds__(BSA) k_adj_nusselt(Dev d, double *fbar, double *theta, double *thf, double R_i,
                                                     double R_o, double wc, double wh)
{
    const int NT = d.NT, NR = d.NR, NZ = d.NZ, NB = d.NB, n = NT * NZ;
    const double *W = d.wall;
    double a1 = 0., a2 = 0.;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a2 += W[(W_NU + 4) * NB + 2 * NT * NR + q];
        a1 += W[(W_NU + 4) * NB + 2 * NT * NR + NT * NZ + q];
    }
    __shared__ double sh[BSA / 64][2];
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[wv][0] = a1; sh[wv][1] = a2; }
    __syncthreads();
    a1 = 0.; a2 = 0.;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) { a1 += sh[v][0]; a2 += sh[v][1]; }
    double *LT = fbar + (size_t)F_T * d.NF;
    double th = 0.;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % NT, k = q / NT;
        // hexagon: Nu_hex = (R_i / a2) sum area dd, dd = -S T_w + A T(0) - B T(1) + C T(2)
        int b = 2 * NT * NR + q;
        double s = wh * R_i / a2 * W[(W_NU + 4) * NB + b];
        th += -s * W[(W_NU + 0) * NB + b];
        thf[q] += -s * W[(W_NU + 0) * NB + b];
        aadd(LT + fidx(d, 2, i, 0, k), s * W[(W_NU + 1) * NB + b]);
        aadd(LT + fidx(d, 2, i, 1, k), -s * W[(W_NU + 2) * NB + b]);
        aadd(LT + fidx(d, 2, i, 2, k), s * W[(W_NU + 3) * NB + b]);
        // cylinder: Nu_cyl = -(R_o / a1) sum area dd (the cold wall value is
        // not a parameter here)
        b = 2 * NT * NR + NT * NZ + q;
        s = -wc * R_o / a1 * W[(W_NU + 4) * NB + b];
        aadd(LT + fidx(d, 1, i, NR - 1, k), s * W[(W_NU + 1) * NB + b]);
        aadd(LT + fidx(d, 1, i, NR - 2, k), -s * W[(W_NU + 2) * NB + b]);
        aadd(LT + fidx(d, 1, i, NR - 3, k), s * W[(W_NU + 3) * NB + b]);
    }
    par_commit(theta + P_THOT, th);
}


// ---------------------------------------------------------------------
// restarted GMRES on (I - G_x^T) lam = seed (the reversed pass as operator)
// ---------------------------------------------------------------------
// Dirichlet slot of the adjoint vector [4N cells | 5NF faces]
__device__ __forceinline__ bool lam_fixed(const Dev &d, long s)
{
    const long c4 = 4 * (long)d.N;
    if (s < c4) return false;
    const long f = s - c4;
    return dirichlet(d, (int)(f / d.NF), (int)(f % d.NF));
}

// out = v - lam (lam = G_x^T v from a reversed pass), 0 in the Dirichlet
// slots, which the pass fills with parameter adjoints
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_apply(Dev d, const VT *v, const double *lam, double *out, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) out[s] = lam_fixed(d, s) ? 0. : (double)v[s] - lam[s];
}

// lam = v (a basis vector, widened to FP64)
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_load(double *lam, const VT *v, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) lam[s] = (double)v[s];
}

// the sum of acc over the block into *out (thread 0 writes)
__device__ __forceinline__ void block_sum_to(double acc, double *out)
{
    __shared__ double sh[BSA / 64];
    acc = wave_sum(acc);
    __syncthreads();  // sh is reused by consecutive calls
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) r += sh[q];
        *out = r;
    }
}

// partial dot products of w with the basis vectors V[0..j), GM_KB of them
// per block row: block (x, g) sums its grid-stride share of V[k] . w for
// k in [g GM_KB, (g + 1) GM_KB) into part[k * gridDim.x + x], so that w is
// read once per GM_KB basis vectors. Products and sums in FP64.
constexpr int GM_KB = 8;
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_dots(const VT *V, long ld, const double *w, long n, int j, double *part)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    const int k0 = blockIdx.y * GM_KB, kn = min(GM_KB, j - k0);
    const VT *v = V + (size_t)k0 * ld;
    double acc[GM_KB];
#pragma unroll
    for (int q = 0; q < GM_KB; q++) acc[q] = 0.;
    for (long s = t; s < n; s += stride) {
        const double ws = w[s];
#pragma unroll
        for (int q = 0; q < GM_KB; q++)
            if (q < kn) acc[q] += (double)v[(size_t)q * ld + s] * ws;
    }
#pragma unroll
    for (int q = 0; q < GM_KB; q++)
        if (q < kn) block_sum_to(acc[q], part + (size_t)(k0 + q) * gridDim.x + blockIdx.x);
}

// partial sums of w . w, block x into part[x]
__global__ void __launch_bounds__(BSA) k_gm_norm2(const double *w, long n, double *part)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    double acc = 0.;
    for (long s = t; s < n; s += stride) acc += w[s] * w[s];
    block_sum_to(acc, part + blockIdx.x);
}

// out[k] = sum of part[k * nb .. k * nb + nb), one block per k, fixed order
__global__ void __launch_bounds__(BSA) k_gm_dots_final(const double *part, int nb, double *out)
{
    const double *p = part + (size_t)blockIdx.x * nb;
    double acc = 0.;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) acc += p[q];
    block_sum_to(acc, out + blockIdx.x);
}

// x += sign * sum_{k < j} c[k] V[k]   (c in device memory)
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_combine(double *x, const VT *V, long ld, int j, const double *c,
                                                    double sign, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) {
        double acc = 0.;
        for (int k = 0; k < j; k++) acc += c[k] * (double)V[(size_t)k * ld + s];
        x[s] += sign * acc;
    }
}

// y = x outside the Dirichlet slots, 0 in them
__global__ void __launch_bounds__(BSA) k_gm_mask(Dev d, double *y, const double *x, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) y[s] = lam_fixed(d, s) ? 0. : x[s];
}

// y = a * x (+ b * z when z is given)
__global__ void __launch_bounds__(BSA) k_gm_scale(double *y, const double *x, double a, const double *z, double b, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) y[s] = z ? a * x[s] + b * z[s] : a * x[s];
}

// v = a * (x - y), rounded to the basis type
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_diff_store(VT *v, const double *x, const double *y, double a, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) v[s] = (VT)(a * (x[s] - y[s]));
}

// a basis vector: v = a * x, rounded to the basis type
template <typename VT>
__global__ void __launch_bounds__(BSA) k_gm_store(VT *v, const double *x, double a, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    for (long s = t; s < n; s += stride) v[s] = (VT)(a * x[s]);
}

}  // namespace

// ---- host side --------------------------------------------------------------
struct nim_adj {
    nim_ctx *ctx = nullptr;
    Adj a{};
    double *seed = nullptr, *prev = nullptr;  // [4N + 5NF] each, like lam (= lc, lf)
    double *cbar = nullptr, *tcbar = nullptr, *fbar = nullptr;
    double *theta = nullptr;                  // [NPAR] accumulated over steps
    double *thf = nullptr;                    // [NT NZ] per hexagon face, accumulated over steps
    // GMRES workspace (allocated on first use): basis V [m + 1][nl] and the
    // augmentation pairs Z, AZ [k][nl] each, of gm_vbytes-byte values; then
    // w, x, b, r_old, dz, z, partial dots and combination coefficients in FP64
    void *gm_block = nullptr;
    double *gm_host = nullptr;  // pinned: coefficients in [0, m + 1), dot results in [m + 1, 2 m + 2)
    int gm_m = 0, gm_k = 0;
    int gm_vbytes = 0;
    int gm_basis = 8;           // bytes per basis value asked for (nim_adj_gmres_options)
    int gm_aug = 3;             // augmentation vectors asked for (nim_adj_gmres_options)
    int gm_deg = 8;             // reversed passes per operator application (nim_adj_gmres_options)
    bool gm_warm = false;       // x holds the last reverse step's solution (the next one starts from it)
    unsigned long long *dmax = nullptr;       // [2]
    unsigned long long *h_dmax = nullptr;
    void *block = nullptr;
};

static double bits_to_double(unsigned long long b)
{
    double x;
    memcpy(&x, &b, sizeof x);
    return x;
}

// grids of the grid-stride loops (bookkeeping kernels) and of the one
// thread per cell kernels
static int grid_for(long n) { return (int)std::min<long>((n + BSA - 1) / BSA, 8192); }
static int cell_grid(long n) { return (int)((n + BSA - 1) / BSA); }
// grid of the grid-stride kernels that sum a scalar parameter: eight blocks
// per CU, one atomic each
static int reduce_grid(long n) { return std::min(cell_grid(n), 2048); }

// doubles of the adjoint's device block: lam, seed, prev, lsrc, lTF, lMF,
// lold, ltc, scr, cbar, tcbar, fbar, par, theta, dmax, thf
static size_t adj_doubles(const Dev &d)
{
    const size_t N = d.N, NF = d.NF, nl = 4 * N + 5 * NF;
    return 3 * nl + NSRC * N + 2 * NCF * N + 4 * N + 3 * N + 10 * N + 4 * N + 3 * N + 5 * NF + 2 * NPAR + 8 +
           (size_t)d.NT * d.NZ;
}

extern "C" int32_t nim_adj_create(nim_ctx *ctx, nim_adj **out)
{
    if (!ctx || !out) return nim_internal_fail(NIM_E_ARG, "null argument");
    *out = nullptr;
    const Dev *d = nim_internal_dev(ctx);
    const size_t N = d->N, NF = d->NF, nl = 4 * N + 5 * NF;
    const size_t total = adj_doubles(*d);
    nim_adj *A = new nim_adj();
    A->ctx = ctx;
    if (hipMalloc(&A->block, total * sizeof(double)) != hipSuccess) {
        delete A;
        return nim_internal_fail(NIM_E_HIP, "adjoint allocation failed");
    }
    if (hipHostMalloc((void **)&A->h_dmax, 2 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(A->block);
        delete A;
        return nim_internal_fail(NIM_E_HIP, "adjoint host allocation failed");
    }
    double *p = (double *)A->block;
    auto take = [&](size_t n) { double *q = p; p += n; return q; };
    A->a.lc = take(nl);
    A->a.lf = A->a.lc + 4 * N;
    A->seed = take(nl);
    A->prev = take(nl);
    A->a.lsrc = take(NSRC * N);
    A->a.lTF = take(NCF * N);
    A->a.lMF = take(NCF * N);
    A->a.lold = take(4 * N);
    A->a.ltc = take(3 * N);
    A->a.scr = take(10 * N);
    A->cbar = take(4 * N);
    A->tcbar = take(3 * N);
    A->fbar = take(5 * NF);
    A->a.par = take(NPAR);
    A->theta = take(NPAR);
    A->dmax = (unsigned long long *)take(8);
    A->thf = take((size_t)d->NT * d->NZ);
    hipStream_t s = (hipStream_t)nim_internal_stream(ctx);
    if (hipMemsetAsync(A->block, 0, total * sizeof(double), s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipHostFree(A->h_dmax);
        (void)hipFree(A->block);
        delete A;
        return nim_internal_fail(NIM_E_HIP, "adjoint clear failed");
    }
    *out = A;
    return NIM_OK;
}

extern "C" int32_t nim_adj_destroy(nim_adj *A)
{
    if (!A) return NIM_OK;
    (void)hipStreamSynchronize((hipStream_t)nim_internal_stream(A->ctx));
    (void)hipHostFree(A->h_dmax);
    (void)hipFree(A->block);
    if (A->gm_block) (void)hipFree(A->gm_block);
    if (A->gm_host) (void)hipHostFree(A->gm_host);
    delete A;
    return NIM_OK;
}

static int32_t adj_check(const char *what)
{
    if (hipGetLastError() != hipSuccess) return nim_internal_fail(NIM_E_HIP, what);
    return NIM_OK;
}

// reverse of a coefficient build (energy 0 momentum, 1 energy with the
// adiabatic walls), then the consumed coefficient adjoints cleared
static void reverse_coefs(nim_adj *A, hipStream_t s, int energy, double *par, int pk)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    hipLaunchKernelGGL(k_adj_coef, dim3(reduce_grid(d.N)), dim3(BSA), 0, s, d, A->a, energy, par, pk);
    hipLaunchKernelGGL(k_adj_coef_clear, dim3(cell_grid(d.N)), dim3(BSA), 0, s, d, A->a, energy);
}

// reverse of a source stage (0 momentum, 1 pressure, 2 energy): the face
// adjoints gathered per face, then the cell terms and the consumed source
// adjoints cleared
static void reverse_sources(nim_adj *A, hipStream_t s, int which)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    const int gN = cell_grid(d.N), gF = cell_grid(d.NF), gR = reduce_grid(d.NF);
    if (which == 0) {
        hipLaunchKernelGGL(k_adj_mom_src_faces, dim3(gR), dim3(BSA), 0, s, d, A->a);
        hipLaunchKernelGGL(k_adj_mom_src_cells, dim3(gN), dim3(BSA), 0, s, d, A->a);
    } else if (which == 1) {
        hipLaunchKernelGGL(k_adj_ppe_factors, dim3(gN), dim3(BSA), 0, s, d, A->a);
        hipLaunchKernelGGL(k_adj_ppe_src_faces, dim3(gF), dim3(BSA), 0, s, d, A->a);
        hipLaunchKernelGGL(k_adj_clear_src, dim3(gN), dim3(BSA), 0, s, d, A->a, (int)R_P);
    } else {
        hipLaunchKernelGGL(k_adj_T_src_faces, dim3(gR), dim3(BSA), 0, s, d, A->a);
        hipLaunchKernelGGL(k_adj_clear_src, dim3(gN), dim3(BSA), 0, s, d, A->a, (int)R_T);
    }
}

// one reverse pass at the context's state: lam <- G_x^T lam, parameter
// adjoints of the pass in lMF, lold, the Dirichlet slots of lf and par
static void reverse_pass(nim_adj *A, hipStream_t s)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    const int per_color = (d.NT / 2) * ((d.NR + 1) / 2) * d.NZ;
    const int gC = (per_color + BSA - 1) / BSA;
    hipLaunchKernelGGL(k_adj_begin, dim3(grid_for(std::max<long>(NCF * (long)d.N, 5 * (long)d.NF))), dim3(BSA), 0, s, d,
                       A->a);
    for (int col = 3; col >= 0; col--) hipLaunchKernelGGL(k_adj_T_sweep, dim3(gC), dim3(BSA), 0, s, d, A->a, col);
    reverse_sources(A, s, 2);
    reverse_coefs(A, s, 1, A->a.par, P_ALPHA);
    for (int col = 3; col >= 0; col--) hipLaunchKernelGGL(k_adj_ppe_sweep, dim3(gC), dim3(BSA), 0, s, d, A->a, col);
    reverse_sources(A, s, 1);
    for (int col = 3; col >= 0; col--)
        hipLaunchKernelGGL(k_adj_mom_sweep, dim3((per_color + 63) / 64), dim3(64, 3), 0, s, d, A->a, col);
    reverse_sources(A, s, 0);
}

extern "C" int32_t nim_adj_seed_nusselt(nim_adj *A, const double weights[3])
{
    if (!A || !weights) return nim_internal_fail(NIM_E_ARG, "null argument");
    const Dev &d = *nim_internal_dev(A->ctx);
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    double R_i, R_o;
    nim_internal_lengths(A->ctx, &R_i, &R_o);
    // out[2] = (out[0] + out[1]) / 2
    const double wc = weights[0] + 0.5 * weights[2], wh = weights[1] + 0.5 * weights[2];
    hipLaunchKernelGGL(k_adj_nusselt, dim3(1), dim3(BSA), 0, s, d, A->fbar, A->theta, A->thf, R_i, R_o, wc, wh);
    return adj_check("adjoint Nusselt seed launch failed");
}

extern "C" int32_t nim_adj_reverse_pass(nim_adj *A)
{
    if (!A) return nim_internal_fail(NIM_E_ARG, "null argument");
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    reverse_pass(A, s);
    int32_t rc = adj_check("adjoint pass launch failed");
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return nim_internal_fail(NIM_E_HIP, "adjoint pass failed");
    return NIM_OK;
}

// end of a reverse step, after the last reversed pass: the hexagon wall
// temperatures and scalar parameters of that pass into the gradient, the
// momentum coefficients' adjoint to the time-centred velocity, and the
// adjoint of the state before the step
static int32_t finish_step(nim_adj *A, hipStream_t s)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    const long nl = 4 * (long)d.N + 5 * (long)d.NF;
    hipLaunchKernelGGL(k_adj_hex, dim3(cell_grid((long)d.NT * d.NZ)), dim3(BSA), 0, s, d, A->a, A->theta, A->thf);
    hipLaunchKernelGGL(k_adj_par_add, dim3(1), dim3(64), 0, s, A->theta, A->a.par);
    if (hipMemsetAsync(A->a.ltc, 0, 3 * (size_t)d.N * sizeof(double), s) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "adjoint clear failed");
    reverse_coefs(A, s, 0, A->theta, P_NU);
    hipLaunchKernelGGL(k_adj_carry, dim3(grid_for(nl)), dim3(BSA), 0, s, d, A->a, A->cbar, A->tcbar, A->fbar);
    int32_t rc = adj_check("adjoint step launch failed");
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return nim_internal_fail(NIM_E_HIP, "adjoint step failed");
    return NIM_OK;
}

extern "C" int32_t nim_adj_reverse_step(nim_adj *A, double tol, int32_t max_passes, int32_t *passes, double *delta)
{
    if (!A || max_passes < 1) return nim_internal_fail(NIM_E_ARG, "bad reverse-step arguments");
    const Dev &d = *nim_internal_dev(A->ctx);
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    const long nl = 4 * (long)d.N + 5 * (long)d.NF;
    const int gL = grid_for(nl);
    // the reversed energy sweep reads the energy coefficients of the
    // converged state, which the forward loop may not have stored
    nim_internal_energy_coefs(A->ctx);
    hipLaunchKernelGGL(k_adj_seed, dim3(gL), dim3(BSA), 0, s, d, A->seed, A->a.lc, A->prev, A->cbar, A->tcbar,
                       A->fbar);
    const int check_every = 10;
    int it = 0;
    double dm = 0., lm = 0.;
    bool conv = false;
    while (it < max_passes) {
        reverse_pass(A, s);
        hipLaunchKernelGGL(k_adj_add_seed, dim3(gL), dim3(BSA), 0, s, A->a.lc, A->seed, nl);
        it++;
        if (it % check_every == 0 || it == max_passes) {
            if (hipMemsetAsync(A->dmax, 0, 2 * sizeof(unsigned long long), s) != hipSuccess)
                return nim_internal_fail(NIM_E_HIP, "adjoint delta clear failed");
            hipLaunchKernelGGL(k_adj_delta, dim3(gL), dim3(BSA), 0, s, A->a.lc, A->prev, nl, A->dmax);
            if (hipMemcpyAsync(A->h_dmax, A->dmax, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return nim_internal_fail(NIM_E_HIP, "adjoint pass failed");
            dm = bits_to_double(A->h_dmax[0]) / check_every;
            lm = bits_to_double(A->h_dmax[1]);
            if (!std::isfinite(dm) || !std::isfinite(lm))
                return nim_internal_fail(NIM_E_NONFINITE, "non-finite adjoint");
            if (dm <= tol * std::max(lm, 1e-300)) {
                conv = true;
                break;
            }
        }
    }
    if (passes) *passes = it;
    if (delta) *delta = lm > 0. ? dm / lm : dm;
    int32_t rc = finish_step(A, s);
    if (rc) return rc;
    return conv ? NIM_OK : nim_internal_fail(NIM_E_MAXITER, "adjoint fixed point not converged");
}

// ---- GMRES ------------------------------------------------------------------
static constexpr int GM_DOT_BLOCKS = 512;

static int32_t gm_alloc(nim_adj *A, int m, int k, int vbytes)
{
    if (A->gm_block && A->gm_m == m && A->gm_k == k && A->gm_vbytes == vbytes) return NIM_OK;
    if (A->gm_block) { (void)hipFree(A->gm_block); A->gm_block = nullptr; }
    if (A->gm_host) { (void)hipHostFree(A->gm_host); A->gm_host = nullptr; }
    if (hipHostMalloc((void **)&A->gm_host, 2 * (size_t)(m + 1) * sizeof(double), 0) != hipSuccess) {
        A->gm_host = nullptr;
        return nim_internal_fail(NIM_E_HIP, "GMRES host buffer allocation failed");
    }
    const Dev &d = *nim_internal_dev(A->ctx);
    const size_t nl = 4 * (size_t)d.N + 5 * (size_t)d.NF;
    const size_t vb = ((size_t)(m + 1 + 2 * k) * nl * vbytes + 255) / 256 * 256;
    const size_t doubles = 6 * nl + (size_t)(m + 1) * GM_DOT_BLOCKS + 2 * (size_t)(m + 1);
    if (hipMalloc(&A->gm_block, vb + doubles * sizeof(double)) != hipSuccess) {
        A->gm_block = nullptr;
        return nim_internal_fail(NIM_E_HIP, "GMRES workspace allocation failed");
    }
    A->gm_m = m;
    A->gm_k = k;
    A->gm_vbytes = vbytes;
    A->gm_warm = false;
    return NIM_OK;
}

// Restarted GMRES on (I - G_x^T) lam = seed, the basis stored as VT, with
// the approximation space of each cycle augmented by the corrections of up
// to k earlier cycles (LGMRES: A. H. Baker, E. R. Jessup and T. Manteuffel,
// SIAM J. Matrix Anal. Appl. 26 (2005) 962-984). A cycle of dimension m takes
// m - n_aug Krylov vectors (one reversed pass each) and then the n_aug stored
// corrections z, whose images A z need no pass: A z = r_old - r_new, the
// residuals before and after the cycle that made z. The flexible Arnoldi
// relation A [V | Z] = V H keeps H upper Hessenberg, so the Givens
// least-squares update is GMRES's. Restarted GMRES stagnates here (the first
// cycle takes the residual down by 1e6, then each restart only halves it);
// the corrections carry the slowly converging directions across restarts.
// Everything but the basis storage is FP64: the operator (a reversed pass
// of the widened vector), the Gram-Schmidt products and sums (classical,
// twice), the iterate and the residual b - A x that starts each cycle and
// decides convergence. An FP32 basis halves the orthogonalisation traffic;
// its rounding (6e-8 relative per stored vector) bounds how far one cycle
// can reduce the residual.
template <typename VT>
static int32_t gmres_solve(nim_adj *A, double tol, int32_t max_passes, int m, int kaug, int pdeg, int32_t *passes,
                           double *relres)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    const long nl = 4 * (long)d.N + 5 * (long)d.NF;
    const int gL = grid_for(nl);
    int32_t rc = gm_alloc(A, m, kaug, (int)sizeof(VT));
    if (rc) return rc;
    VT *V = (VT *)A->gm_block, *Za = V + (size_t)(m + 1) * nl, *AZa = Za + (size_t)kaug * nl;
    double *w = (double *)((char *)A->gm_block + ((size_t)(m + 1 + 2 * kaug) * nl * sizeof(VT) + 255) / 256 * 256);
    double *x = w + nl, *b = x + nl, *rold = b + nl, *dz = rold + nl, *zb = dz + nl;
    double *part = zb + nl, *dh = part + (size_t)(m + 1) * GM_DOT_BLOCKS, *dc = dh + (m + 1);
    double *lam = A->a.lc, *hin = A->gm_host, *hout = A->gm_host + (m + 1);
    const int nbd = (int)std::min<long>(GM_DOT_BLOCKS, (nl + BSA - 1) / BSA);
    nim_internal_energy_coefs(A->ctx);
    // the seed (b), as the plain iteration starts from it
    hipLaunchKernelGGL(k_adj_seed, dim3(gL), dim3(BSA), 0, s, d, A->seed, lam, A->prev, A->cbar, A->tcbar, A->fbar);
    // dots of vec with V[0..j) into host h
    auto dots = [&](const double *vec, int j, double *h) -> int32_t {
        hipLaunchKernelGGL(k_gm_dots<VT>, dim3(nbd, (j + GM_KB - 1) / GM_KB), dim3(BSA), 0, s, (const VT *)V, nl, vec,
                           nl, j, part);
        hipLaunchKernelGGL(k_gm_dots_final, dim3(j), dim3(BSA), 0, s, (const double *)part, nbd, dh);
        if (hipMemcpyAsync(hout, dh, j * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return nim_internal_fail(NIM_E_HIP, "GMRES dot products failed");
        std::copy(hout, hout + j, h);
        return NIM_OK;
    };
    auto norm_of = [&](const double *vec, double *out) -> int32_t {
        hipLaunchKernelGGL(k_gm_norm2, dim3(nbd), dim3(BSA), 0, s, vec, nl, part);
        hipLaunchKernelGGL(k_gm_dots_final, dim3(1), dim3(BSA), 0, s, (const double *)part, nbd, dh);
        if (hipMemcpyAsync(hout, dh, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return nim_internal_fail(NIM_E_HIP, "GMRES norm failed");
        *out = std::sqrt(std::max(hout[0], 0.));
        return NIM_OK;
    };
    // y (device, n values) from host
    auto upload = [&](const double *y, int n) -> int32_t {
        if (hipStreamSynchronize(s) != hipSuccess) return nim_internal_fail(NIM_E_HIP, "GMRES failed");
        std::copy(y, y + n, hin);
        if (hipMemcpyAsync(dc, hin, n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
            return nim_internal_fail(NIM_E_HIP, "GMRES copy failed");
        return NIM_OK;
    };
    // out = (I - (G_x^T)^deg) v, v a basis vector or x: deg reversed passes,
    // the Dirichlet slots cleared between them
    int ops = 0;
    auto apply = [&](const auto *v, double *out, int deg) {
        using T = std::remove_cv_t<std::remove_pointer_t<decltype(v)>>;
        hipLaunchKernelGGL(k_gm_load<T>, dim3(gL), dim3(BSA), 0, s, lam, v, nl);
        for (int q = 0; q < deg; q++) {
            if (q > 0) hipLaunchKernelGGL(k_gm_mask, dim3(gL), dim3(BSA), 0, s, d, lam, (const double *)lam, nl);
            reverse_pass(A, s);
        }
        hipLaunchKernelGGL(k_gm_apply<T>, dim3(gL), dim3(BSA), 0, s, d, v, (const double *)lam, out, nl);
        ops += deg;
    };
    // the Neumann preconditioner: out = sum_{q < deg} (G_x^T)^q r
    auto precondition = [&](const double *r, double *out, int deg) {
        hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, out, r, 1., (const double *)nullptr, 0., nl);
        hipLaunchKernelGGL(k_gm_load<double>, dim3(gL), dim3(BSA), 0, s, lam, r, nl);
        for (int q = 1; q < deg; q++) {
            reverse_pass(A, s);
            hipLaunchKernelGGL(k_gm_mask, dim3(gL), dim3(BSA), 0, s, d, lam, (const double *)lam, nl);
            hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, out, (const double *)out, 1.,
                               (const double *)lam, 1., nl);
            ops++;
        }
    };
    // b = the seed without its Dirichlet slots (they hold no unknown); x
    // starts from the previous reverse step's solution (consecutive steps'
    // adjoints differ little), or 0
    hipLaunchKernelGGL(k_gm_mask, dim3(gL), dim3(BSA), 0, s, d, b, (const double *)A->seed, nl);
    bool x_zero = !A->gm_warm;
    if (x_zero && hipMemsetAsync(x, 0, nl * sizeof(double), s) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "GMRES clear failed");
    static const bool verbose = [] {
        const char *v = getenv("NIM_VERBOSE");
        return v && atoi(v) != 0;
    }();
    double bnorm = 0.;
    if ((rc = norm_of(b, &bnorm))) return rc;
    double res = bnorm, true_res = bnorm;
    bool conv = bnorm == 0.;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1), y(m), hh(m + 1);
    // augmentation: `filled` pairs in slots [0, filled) (filled in order, then
    // the oldest replaced); the correction of the last cycle waits in slot
    // `pend` for its image, which the next cycle's residual gives before any
    // pair is used
    int filled = 0, pend = -1, next = 0;
    double dz_norm = 0.;
    while (!conv && ops < max_passes) {
        // the true residual r = b - (I - G_x^T) x (in w) decides convergence
        if (x_zero) {
            hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, w, (const double *)b, 1., (const double *)nullptr,
                               0., nl);
        } else {
            apply((const double *)x, w, 1);
            hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, w, (const double *)b, 1., (const double *)w, -1.,
                               nl);
        }
        if ((rc = norm_of(w, &true_res))) return rc;
        if (!std::isfinite(true_res)) return nim_internal_fail(NIM_E_NONFINITE, "non-finite adjoint residual");
        res = true_res;
        if (verbose) {
            fprintf(stderr, "GMRES: %d reversed passes, relative residual %.3e\n", ops,
                    bnorm > 0. ? true_res / bnorm : 0.);
            fflush(stderr);
        }
        if (true_res <= tol * bnorm) { conv = true; break; }
        // the cycle's system: M (I - G_x^T) lam = M b with M the Neumann
        // polynomial of degree pdeg - 1, i.e. (I - (G_x^T)^pdeg) lam = M b;
        // its residual z = M r starts the cycle, V[0] = z / |z|
        double *z = w;
        if (pdeg > 1) {
            precondition(w, zb, pdeg);
            z = zb;
        }
        if (pend >= 0) {
            // the operator's image of the stored correction dz / |dz|:
            // (z_old - z) / |dz|
            hipLaunchKernelGGL(k_gm_diff_store<VT>, dim3(gL), dim3(BSA), 0, s, AZa + (size_t)pend * nl,
                               (const double *)rold, (const double *)z, 1. / dz_norm, nl);
            pend = -1;
        }
        const int naug = filled;
        double beta;
        if (pdeg > 1) {
            if ((rc = norm_of(z, &beta))) return rc;
            if (!std::isfinite(beta)) return nim_internal_fail(NIM_E_NONFINITE, "non-finite adjoint residual");
        } else {
            beta = true_res;
        }
        // the cycle stops where its residual estimate stands to beta as the
        // tolerance does to the true residual
        const double cycle_tol = tol * bnorm * (beta / true_res);
        if (kaug > 0)
            hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, rold, (const double *)z, 1.,
                               (const double *)nullptr, 0., nl);
        hipLaunchKernelGGL(k_gm_store<VT>, dim3(gL), dim3(BSA), 0, s, V, (const double *)z, 1. / beta, nl);
        std::fill(g.begin(), g.end(), 0.);
        g[0] = beta;
        const int nk = m - std::min(naug, m - 1);  // Krylov vectors of this cycle, then the corrections
        int jdone = 0;
        for (int j = 0; j < m && ops < max_passes; j++) {
            if (j < nk)
                apply((const VT *)(V + (size_t)j * nl), w, pdeg);
            else
                hipLaunchKernelGGL(k_gm_load<VT>, dim3(gL), dim3(BSA), 0, s, w,
                                   (const VT *)(AZa + (size_t)(j - nk) * nl), nl);
            // classical Gram-Schmidt, twice
            std::fill(h.begin(), h.end(), 0.);
            for (int pass = 0; pass < 2; pass++) {
                if ((rc = dots(w, j + 1, hh.data()))) return rc;  // (synchronises: hin is free)
                std::copy(hh.begin(), hh.begin() + j + 1, hin);
                if (hipMemcpyAsync(dc, hin, (j + 1) * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
                    return nim_internal_fail(NIM_E_HIP, "GMRES copy failed");
                hipLaunchKernelGGL(k_gm_combine<VT>, dim3(gL), dim3(BSA), 0, s, w, (const VT *)V, nl, j + 1,
                                   (const double *)dc, -1., nl);
                for (int k = 0; k <= j; k++) h[k] += hh[k];
            }
            double hn;
            if ((rc = norm_of(w, &hn))) return rc;
            if (!std::isfinite(hn)) return nim_internal_fail(NIM_E_NONFINITE, "non-finite adjoint Krylov vector");
            h[j + 1] = hn;
            if (hn > 0.)
                hipLaunchKernelGGL(k_gm_store<VT>, dim3(gL), dim3(BSA), 0, s, V + (size_t)(j + 1) * nl,
                                   (const double *)w, 1. / hn, nl);
            // Givens rotations of the new column
            for (int k = 0; k < j; k++) {
                const double t = cs[k] * h[k] + sn[k] * h[k + 1];
                h[k + 1] = -sn[k] * h[k] + cs[k] * h[k + 1];
                h[k] = t;
            }
            const double r = std::hypot(h[j], h[j + 1]);
            cs[j] = r > 0. ? h[j] / r : 1.;
            sn[j] = r > 0. ? h[j + 1] / r : 0.;
            h[j] = r;
            h[j + 1] = 0.;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = h[k];
            jdone = j + 1;
            res = std::fabs(g[j + 1]);
            if (res <= cycle_tol || hn == 0.) break;
        }
        if (jdone == 0) break;
        // H y = g (upper triangular); dz = [V | Z] y; x += dz
        for (int k = jdone - 1; k >= 0; k--) {
            double t = g[k];
            for (int q = k + 1; q < jdone; q++) t -= H[(size_t)k * m + q] * y[q];
            y[k] = t / H[(size_t)k * m + k];
        }
        if ((rc = upload(y.data(), jdone))) return rc;
        if (hipMemsetAsync(dz, 0, nl * sizeof(double), s) != hipSuccess)
            return nim_internal_fail(NIM_E_HIP, "GMRES clear failed");
        hipLaunchKernelGGL(k_gm_combine<VT>, dim3(gL), dim3(BSA), 0, s, dz, (const VT *)V, nl, std::min(jdone, nk),
                           (const double *)dc, 1., nl);
        if (jdone > nk)
            hipLaunchKernelGGL(k_gm_combine<VT>, dim3(gL), dim3(BSA), 0, s, dz, (const VT *)Za, nl, jdone - nk,
                               (const double *)(dc + nk), 1., nl);
        hipLaunchKernelGGL(k_gm_scale, dim3(gL), dim3(BSA), 0, s, x, (const double *)x, 1., (const double *)dz, 1.,
                           nl);
        x_zero = false;
        // keep dz / |dz| as a correction (its image follows from the next
        // cycle's residual), replacing the oldest when all k are in use
        if (kaug > 0) {
            if ((rc = norm_of(dz, &dz_norm))) return rc;
            if (dz_norm > 0. && std::isfinite(dz_norm)) {
                pend = next;
                next = (next + 1) % kaug;
                filled = std::min(filled + 1, kaug);
                hipLaunchKernelGGL(k_gm_store<VT>, dim3(gL), dim3(BSA), 0, s, Za + (size_t)pend * nl,
                                   (const double *)dz, 1. / dz_norm, nl);
            }
        }
        // the next cycle starts from the true residual of x and stops there
        // if it is within the tolerance
    }
    if (rc = adj_check("GMRES launch failed"); rc) return rc;
    // the parameter adjoints of lam = x: one more reversed pass, then the
    // seed added, as the plain iteration ends
    hipLaunchKernelGGL(k_gm_load<double>, dim3(gL), dim3(BSA), 0, s, lam, (const double *)x, nl);
    reverse_pass(A, s);
    hipLaunchKernelGGL(k_adj_add_seed, dim3(gL), dim3(BSA), 0, s, lam, (const double *)A->seed, nl);
    ops++;
    if (passes) *passes = ops;
    if (relres) *relres = bnorm > 0. ? true_res / bnorm : 0.;
    A->gm_warm = conv;
    if ((rc = finish_step(A, s))) return rc;
    return conv ? NIM_OK : nim_internal_fail(NIM_E_MAXITER, "adjoint GMRES not converged");
}

extern "C" int32_t nim_adj_reverse_step_gmres(nim_adj *A, double tol, int32_t max_passes, int32_t restart,
                                              int32_t *passes, double *relres)
{
    if (!A || max_passes < 1 || restart < 1 || restart > 200 || !(tol > 0.))
        return nim_internal_fail(NIM_E_ARG, "bad GMRES reverse-step arguments");
    const int k = std::min(A->gm_aug, restart - 1);
    return A->gm_basis == 8 ? gmres_solve<double>(A, tol, max_passes, restart, k, A->gm_deg, passes, relres)
                            : gmres_solve<float>(A, tol, max_passes, restart, k, A->gm_deg, passes, relres);
}

extern "C" int32_t nim_adj_gmres_options(nim_adj *A, int32_t basis_bytes, int32_t augment, int32_t degree)
{
    if (!A || (basis_bytes != 4 && basis_bytes != 8) || augment < 0 || augment > 16 || degree < 1 || degree > 64)
        return nim_internal_fail(NIM_E_ARG, "GMRES options: basis of 4 or 8 bytes, 0 to 16 augmentation vectors, "
                                            "1 to 64 passes per operator application");
    A->gm_basis = basis_bytes;
    A->gm_aug = augment;
    A->gm_deg = degree;
    return NIM_OK;
}

extern "C" int32_t nim_adj_gradient_faces(nim_adj *A, double *out, int64_t count)
{
    if (!A || !out) return nim_internal_fail(NIM_E_ARG, "null argument");
    const Dev &d = *nim_internal_dev(A->ctx);
    if (count != (int64_t)d.NT * d.NZ)
        return nim_internal_fail(NIM_E_ARG, "nim_adj_gradient_faces: count must be n_theta * n_z");
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    if (hipMemcpyAsync(out, A->thf, count * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "adjoint gradient copy failed");
    return NIM_OK;
}

extern "C" int32_t nim_adj_gradient(nim_adj *A, double out[4])
{
    if (!A || !out) return nim_internal_fail(NIM_E_ARG, "null argument");
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    if (hipMemcpyAsync(out, A->theta, NPAR * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "adjoint gradient copy failed");
    return NIM_OK;
}

extern "C" int32_t nim_adj_clear(nim_adj *A)
{
    if (!A) return nim_internal_fail(NIM_E_ARG, "null argument");
    const Dev &d = *nim_internal_dev(A->ctx);
    const size_t total = adj_doubles(d);
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    A->gm_warm = false;
    if (hipMemsetAsync(A->block, 0, total * sizeof(double), s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "adjoint clear failed");
    return NIM_OK;
}

// the reverse of one forward stage (nim_run_stage numbering: 0 momentum
// sources, 1 momentum sweep, 2 pressure sources, 3 pressure sweep, 4 energy
// coefficients and adiabatic walls, 5 energy sources, 6 energy sweep) or,
// as stage 7, of the momentum coefficient build; scalar parameter adjoints
// go to the "parameters" buffer
extern "C" int32_t nim_adj_reverse_stage(nim_adj *A, int32_t stage)
{
    if (!A || stage < 0 || stage > 7) return nim_internal_fail(NIM_E_ARG, "bad reverse-stage arguments");
    const Dev &d = *nim_internal_dev(A->ctx);
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    const int per_color = (d.NT / 2) * ((d.NR + 1) / 2) * d.NZ;
    const int gC = (per_color + BSA - 1) / BSA;
    switch (stage) {
    case 0: reverse_sources(A, s, 0); break;
    case 1:
        for (int col = 3; col >= 0; col--)
            hipLaunchKernelGGL(k_adj_mom_sweep, dim3((per_color + 63) / 64), dim3(64, 3), 0, s, d, A->a, col);
        break;
    case 2: reverse_sources(A, s, 1); break;
    case 3:
        for (int col = 3; col >= 0; col--) hipLaunchKernelGGL(k_adj_ppe_sweep, dim3(gC), dim3(BSA), 0, s, d, A->a, col);
        break;
    case 4: reverse_coefs(A, s, 1, A->a.par, P_ALPHA); break;
    case 5: reverse_sources(A, s, 2); break;
    case 6:
        for (int col = 3; col >= 0; col--) hipLaunchKernelGGL(k_adj_T_sweep, dim3(gC), dim3(BSA), 0, s, d, A->a, col);
        break;
    default: reverse_coefs(A, s, 0, A->a.par, P_NU); break;
    }
    int32_t rc = adj_check("adjoint stage launch failed");
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return nim_internal_fail(NIM_E_HIP, "adjoint stage failed");
    return NIM_OK;
}

// the adjoint of a context array, by the context array's name and in the
// reference layout (nim_get_array / nim_set_array), or the scalar
// "parameters" (this pass) and "gradient" (accumulated) in the order
// nu, Rinv, alpha, T_hot
static bool adj_lookup(nim_adj *A, const char *name, double *&p, int &kind, int &mult)
{
    const Dev &d = *nim_internal_dev(A->ctx);
    const size_t N = d.N, NF = d.NF;
    std::string s(name);
    const char *ff[5] = {"faces_u", "faces_v", "faces_w", "faces_p", "faces_T"};
    for (int q = 0; q < 5; q++)
        if (s == ff[q]) { p = A->a.lf + q * NF; kind = 1; mult = 1; return true; }
    const char *cf[4] = {"u", "v", "w", "T"};
    for (int q = 0; q < 4; q++) {
        if (s == std::string(cf[q]) + "xyz") { p = A->a.lc + q * N; kind = 0; mult = 1; return true; }
        if (s == std::string(cf[q]) + "xyz_old") { p = A->a.lold + q * N; kind = 0; mult = 1; return true; }
        if (q < 3 && s == std::string(cf[q]) + "xyzt") { p = A->a.ltc + q * N; kind = 0; mult = 1; return true; }
    }
    const char *src[5] = {"R_B_u", "R_B_v", "R_B_w", "R_b_p", "R_D_T"};
    for (int q = 0; q < 5; q++)
        if (s == src[q]) { p = A->a.lsrc + q * N; kind = 0; mult = 1; return true; }
    const char *cn[4] = {"F4", "F1", "F2", "F3"};
    const int co[4] = {C_F4, C_F1, C_F2, C_F3}, cm[4] = {8, 5, 5, 5};
    for (int q = 0; q < 4; q++) {
        if (s == std::string("M_") + cn[q]) { p = A->a.lMF + co[q] * N; kind = 0; mult = cm[q]; return true; }
        if (s == std::string("MT_") + cn[q]) { p = A->a.lTF + co[q] * N; kind = 0; mult = cm[q]; return true; }
    }
    if (s == "parameters") { p = A->a.par; kind = -1; mult = NPAR; return true; }
    if (s == "gradient") { p = A->theta; kind = -1; mult = NPAR; return true; }
    return false;
}

static int32_t adj_move(nim_adj *A, const char *name, double *host, int64_t count, bool to_host)
{
    if (!A || !name || !host) return nim_internal_fail(NIM_E_ARG, "null argument");
    double *p;
    int kind, mult;
    if (!adj_lookup(A, name, p, kind, mult)) return nim_internal_fail(NIM_E_ARG, "unknown adjoint array");
    hipStream_t s = (hipStream_t)nim_internal_stream(A->ctx);
    if (hipStreamSynchronize(s) != hipSuccess) return nim_internal_fail(NIM_E_HIP, "adjoint copy failed");
    if (kind < 0) {
        if (count != mult) return nim_internal_fail(NIM_E_ARG, "wrong element count");
        hipError_t e = to_host ? hipMemcpy(host, p, mult * sizeof(double), hipMemcpyDeviceToHost)
                               : hipMemcpy(p, host, mult * sizeof(double), hipMemcpyHostToDevice);
        return e == hipSuccess ? NIM_OK : nim_internal_fail(NIM_E_HIP, "adjoint copy failed");
    }
    return nim_internal_move(A->ctx, p, kind, mult, name, host, count, to_host);
}

extern "C" int32_t nim_adj_get(nim_adj *A, const char *name, double *out, int64_t count)
{
    return adj_move(A, name, out, count, true);
}

extern "C" int32_t nim_adj_set(nim_adj *A, const char *name, const double *in, int64_t count)
{
    return adj_move(A, name, const_cast<double *>(in), count, false);
}
