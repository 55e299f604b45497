// HIP kernels of the NIM time step for gfx950 (FP64).
//
// One thread handles one cell (or one wall face) and performs the same
// per-cell arithmetic as the reference subroutine named in each comment.
// Relaxation sweeps run in four colour phases, colour = (2 i + j + 3 k) mod 4:
// within a colour no cell reads a value another cell of that colour writes
// (the stencils reach +-1 in every direction, +-2 in r and z at walls, and
// the diagonal neighbours of the face-tangential corrections), so a phase is
// race-free and needs no atomics. Convergence increments are reduced inside
// the sweeps: each value is written once per outer iteration, so
// |new - old| at the write is exactly the find_error increment
// (nse.f90:260-325); waves reduce with shuffles and one lane per wave does an
// atomic max on the bit pattern (non-negative doubles order like uint64).
#include "nim_stencil.h"

namespace nim {

__device__ __forceinline__ void err_commit(const Dev &d, int field, double e)
{
    e = e <= DBL_MAX ? e : INFINITY;  // fmax drops NaN: carry it as infinity
    e = wave_max(e);
    if ((threadIdx.x & 63) == 0) {
        int slot = d.ctl->iter & 1;
        unsigned long long *p = d.errs + err_slot(slot, field, blockIdx.x & (ERR_COPIES - 1));
        atomicMax(p, (unsigned long long)__double_as_longlong(e));
    }
}

// key of a signed double that orders like the double under unsigned
// comparison (positive: sign bit set; negative: all bits flipped); key 0 is
// below every key of a number and marks "no value"
__device__ __host__ __forceinline__ unsigned long long order_key(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : b | 0x8000000000000000ull;
}
__device__ __forceinline__ double order_value(unsigned long long k)
{
    return __longlong_as_double((long long)((k >> 63) ? k & 0x7fffffffffffffffull : ~k));
}

// the pressure increments of a pass: max |dp|, max dp and max -dp
struct PErr {
    double a = 0., hi = -INFINITY, lo = -INFINITY;
    __device__ void add(double dp)
    {
        a = err_max(a, dp);
        hi = fmax(hi, dp);
        lo = fmax(lo, -dp);
    }
};

__device__ __forceinline__ void err_commit_p(const Dev &d, const PErr &e)
{
    err_commit(d, F_P, e.a);
    // (the extra reductions and atomics only when the solve asked for the
    // half range: uniform over the launch)
    if (d.ctl->p_measure != P_HALFRANGE) return;
    const double hi = wave_max(e.hi), lo = wave_max(e.lo);
    if ((threadIdx.x & 63) == 0 && hi > -INFINITY) {
        const int slot = d.ctl->iter & 1, cp = blockIdx.x & (ERR_COPIES - 1);
        atomicMax(d.errs + err_slot(slot, E_PHI, cp), order_key(hi));
        atomicMax(d.errs + err_slot(slot, E_PLO, cp), order_key(lo));
    }
}

// cell of this thread in one launch of a sweep schedule: the item that owns
// the block, then the colour cell within the item
__device__ __forceinline__ bool sched_cell(const Dev &d, const Sched &s, int lb, int &i, int &j, int &k)
{
    if (s.rsb) {
        // XCD x takes rows-of-this-colour [x*nrh/8, (x+1)*nrh/8) of every
        // plane (a radial slab), plane after plane (from the top with rev)
        const int nth = d.NT >> 1, nrh = (d.NR + 1) >> 1, q = s.bstart[1] / NXCD;
        const int x = lb / q, r = lb - x * q, kk = r / s.rsb, kp = s.rev ? d.NZ - 1 - kk : kk;
        const int tid = kp * (nth * nrh) + x * (nth * (nrh / NXCD)) + (r - kk * s.rsb) * (int)blockDim.x +
                        (int)threadIdx.x;
        return color_cell(d, s.color[0], s.k0[0], s.k1[0], tid, i, j, k);
    }
    int q = 0;
#pragma unroll
    for (int t = 1; t < MAX_ITEMS; t++)
        if (t < s.n && lb >= s.bstart[t]) q = t;
    int tid = (lb - s.bstart[q]) * (blockDim.x) + threadIdx.x;
    return color_cell(d, s.color[q], s.k0[q], s.k1[q], tid, i, j, k);
}

// exponential-fit constants of every cell: momentum (time-centred velocity,
// nu; nim.f90:166-290) or energy (mean of current iterate and previous
// level, alpha; nim.f90:339-463)
__device__ __forceinline__ void coef_A_cell(const Dev &d, int energy, int c)
{
    const int N = d.N;
    const double *g = d.gc;
    double ux, uy, uz, kd;
    if (energy) {
        ux = 0.5 * (d.sc[(S_CUR + 0) * N + c] + d.sc[(S_OLD + 0) * N + c]);
        uy = 0.5 * (d.sc[(S_CUR + 1) * N + c] + d.sc[(S_OLD + 1) * N + c]);
        uz = 0.5 * (d.sc[(S_CUR + 2) * N + c] + d.sc[(S_OLD + 2) * N + c]);
        kd = d.alpha;
    } else {
        ux = d.sc[(S_TC + 0) * N + c];
        uy = d.sc[(S_TC + 1) * N + c];
        uz = d.sc[(S_TC + 2) * N + c];
        kd = d.nu;
    }
    double *out = energy ? d.TA : d.MA;
    for (int dir = 0; dir < 3; dir++) {
        double vel = ux * g[(G_XI + 3 * dir + 0) * N + c] + uy * g[(G_XI + 3 * dir + 1) * N + c]
                     + uz * g[(G_XI + 3 * dir + 2) * N + c];
        double A[6];
        nim_direction(vel, g[(G_GD + dir) * N + c] * kd, g[(G_GF + 2 * dir + 1) * N + c],
                      g[(G_GF + 2 * dir) * N + c], d.v_limit, A);
        for (int m = 0; m < 6; m++) out[(dir * 6 + m) * N + c] = A[m];
    }
}

__global__ void __launch_bounds__(256) k_coef_A(Dev d, int energy)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (energy && d.ctl->done) return;
    if (c >= d.N) return;
    coef_A_cell(d, energy, c);
}

// F4, F1..F3 from the constants (nim.f90:292-329 / 465-502); for the energy
// equation the same kernel also applies the adiabatic top/bottom wall
// extrapolation apply_zero_flux (bc.f90:345-390), which reads only interior
// axial faces
__device__ __forceinline__ void coef_F_cell(const Dev &d, int energy, int c, double &eT)
{
    if (c < d.N) {
        const int N = d.N;
        const double *A = energy ? d.TA : d.MA;
        double *F = energy ? d.TF : d.MF;
        int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
#define AV(n, dir, m) A[((dir - 1) * 6 + (m) - 1) * N + (n)]
        double f2 = 0.5 * (1. / AV(c, 1, 5) + 1. / AV(c, 2, 5) + 1. / AV(c, 3, 5) - 1. / d.DT);
        double f3 = -(AV(c, 1, 6) / AV(c, 1, 5)), f4 = (AV(c, 1, 6) - 1.) / AV(c, 1, 5);
        double f5 = -(AV(c, 2, 6) / AV(c, 2, 5)), f6 = (AV(c, 2, 6) - 1.) / AV(c, 2, 5);
        double f7 = -(AV(c, 3, 6) / AV(c, 3, 5)), f8 = (AV(c, 3, 6) - 1.) / AV(c, 3, 5);
        F[(C_F4 + 0) * N + c] = f2 + f3 + f4 + f5 + f6 + f7 + f8;
        F[(C_F4 + 1) * N + c] = f2;
        F[(C_F4 + 2) * N + c] = f3;
        F[(C_F4 + 3) * N + c] = f4;
        F[(C_F4 + 4) * N + c] = f5;
        F[(C_F4 + 5) * N + c] = f6;
        F[(C_F4 + 6) * N + c] = f7;
        F[(C_F4 + 7) * N + c] = f8;
        const int base[3] = {C_F1, C_F2, C_F3};
        for (int dir = 1; dir <= 3; dir++) {
            int c2;
            if (dir == 1) { if (j == d.NR - 1) continue; c2 = c + d.NT; }
            else if (dir == 2) c2 = cidx(d, ip1(d, i), j, k);
            else { if (k == d.NZ - 1) continue; c2 = c + d.NT * d.NR; }
            double b = AV(c, dir, 3) - (AV(c, dir, 1) * AV(c, dir, 6) / AV(c, dir, 5));
            double e = AV(c2, dir, 4) - AV(c2, dir, 2) / AV(c2, dir, 5) + AV(c2, dir, 2) * AV(c2, dir, 6) / AV(c2, dir, 5);
            double h = AV(c, dir, 1) / (2. * AV(c, dir, 5));
            double q = AV(c2, dir, 2) / (2. * AV(c2, dir, 5));
            int s = base[dir - 1];
            F[(s + 0) * N + c] = b + e + 2. * h + 2. * q;
            F[(s + 1) * N + c] = b;
            F[(s + 2) * N + c] = e;
            F[(s + 3) * N + c] = h;
            F[(s + 4) * N + c] = q;
        }
#undef AV
        if (energy && (k == 0 || k == d.NZ - 1)) {
            double *T = d.fc + (size_t)F_T * d.NF;
            int b, fw, fa, fb, fcc;
            if (k == 0) {
                b = i + d.NT * j;
                fw = fidx(d, 5, i, j, 0); fa = fidx(d, 6, i, j, 0); fb = fidx(d, 6, i, j, 1); fcc = fidx(d, 6, i, j, 2);
            } else {
                b = d.NT * d.NR + i + d.NT * j;
                fw = fidx(d, 6, i, j, k); fa = fidx(d, 5, i, j, k); fb = fidx(d, 5, i, j, k - 1); fcc = fidx(d, 5, i, j, k - 2);
            }
            const double *W = d.wall;
            const int NB = d.NB;
            double nv = (W[(W_ZF + 0) * NB + b] * T[fa] - W[(W_ZF + 1) * NB + b] * T[fb] + W[(W_ZF + 2) * NB + b] * T[fcc])
                        / W[(W_ZF + 3) * NB + b];
            eT = fabs(nv - T[fw]);
            T[fw] = nv;
        }
    }
}

__global__ void __launch_bounds__(256) k_coef_F(Dev d, int energy)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (energy && d.ctl->done) return;
    double eT = 0.;
    coef_F_cell(d, energy, c, eT);
    if (energy) err_commit(d, F_T, eT);
}

// Energy coefficients of one outer iteration in a single pass
// (calculate_temperature_nim_constants_formulation2, nim.f90:334-503, then
// apply_zero_flux, bc.f90:345-390): each cell evaluates its own three
// directions and, for the face coefficients F1..F3, the one direction it
// needs of each + neighbour, so the 18 exponential-fit constants per cell
// never leave registers.
__device__ __forceinline__ void energy_direction(const Dev &d, int c, int dir, double A[6])
{
    const int N = d.N;
    const double *g = d.gc;
    double ux = 0.5 * (d.sc[(S_CUR + 0) * N + c] + d.sc[(S_OLD + 0) * N + c]);
    double uy = 0.5 * (d.sc[(S_CUR + 1) * N + c] + d.sc[(S_OLD + 1) * N + c]);
    double uz = 0.5 * (d.sc[(S_CUR + 2) * N + c] + d.sc[(S_OLD + 2) * N + c]);
    double vel = ux * g[(G_XI + 3 * dir + 0) * N + c] + uy * g[(G_XI + 3 * dir + 1) * N + c]
                 + uz * g[(G_XI + 3 * dir + 2) * N + c];
    nim_direction(vel, g[(G_GD + dir) * N + c] * d.alpha, g[(G_GF + 2 * dir + 1) * N + c], g[(G_GF + 2 * dir) * N + c],
                  d.v_limit, A);
}

// The energy coefficients of cell c = (i, j, k), handed to put(slot, value)
// in the slots of Dev::TF (nim.f90:465-502): F4 always, F1..F3 where the +
// neighbour in that direction exists
template <class Put>
__device__ __forceinline__ void energy_coefs(const Dev &d, int c, int i, int j, int k, Put put)
{
    double A[3][6];
#pragma unroll
    for (int dir = 0; dir < 3; dir++) energy_direction(d, c, dir, A[dir]);
    double f2 = 0.5 * (1. / A[0][4] + 1. / A[1][4] + 1. / A[2][4] - 1. / d.DT);
    double f3 = -(A[0][5] / A[0][4]), f4 = (A[0][5] - 1.) / A[0][4];
    double f5 = -(A[1][5] / A[1][4]), f6 = (A[1][5] - 1.) / A[1][4];
    double f7 = -(A[2][5] / A[2][4]), f8 = (A[2][5] - 1.) / A[2][4];
    put(C_F4 + 0, f2 + f3 + f4 + f5 + f6 + f7 + f8);
    put(C_F4 + 1, f2);
    put(C_F4 + 2, f3);
    put(C_F4 + 3, f4);
    put(C_F4 + 4, f5);
    put(C_F4 + 5, f6);
    put(C_F4 + 6, f7);
    put(C_F4 + 7, f8);
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        int c2;
        if (dir == 0) { if (j == d.NR - 1) continue; c2 = c + d.NT; }
        else if (dir == 1) c2 = cidx(d, ip1(d, i), j, k);
        else { if (k == d.NZ - 1) continue; c2 = c + d.NT * d.NR; }
        double B[6];
        energy_direction(d, c2, dir, B);
        double b = A[dir][2] - (A[dir][0] * A[dir][5] / A[dir][4]);
        double e = B[3] - B[1] / B[4] + B[1] * B[5] / B[4];
        double h = A[dir][0] / (2. * A[dir][4]);
        double q = B[1] / (2. * B[4]);
        const int s = dir == 0 ? C_F1 : dir == 1 ? C_F2 : C_F3;
        put(s + 0, b + e + 2. * h + 2. * q);
        put(s + 1, b);
        put(s + 2, e);
        put(s + 3, h);
        put(s + 4, q);
    }
}

// apply_zero_flux (bc.f90:345-390) for the adiabatic wall face of cell
// (i, j, k), k = 0 or NZ - 1: relax it (GS) or write its residual (RES)
template <int MODE>
__device__ __forceinline__ void zero_flux_wall(const Dev &d, const View<double> &v, int i, int j, int k, double &eT)
{
    double *T = d.fc + (size_t)F_T * d.NF;
    int b, fw, fa, fb, fcc;
    if (k == 0) {
        b = i + d.NT * j;
        fw = fidx(d, 5, i, j, 0); fa = fidx(d, 6, i, j, 0); fb = fidx(d, 6, i, j, 1); fcc = fidx(d, 6, i, j, 2);
    } else {
        b = d.NT * d.NR + i + d.NT * j;
        fw = fidx(d, 6, i, j, k); fa = fidx(d, 5, i, j, k); fb = fidx(d, 5, i, j, k - 1); fcc = fidx(d, 5, i, j, k - 2);
    }
    const double *W = d.wall;
    const int NB = d.NB;
    double S = W[(W_ZF + 0) * NB + b] * T[fa] - W[(W_ZF + 1) * NB + b] * T[fb] + W[(W_ZF + 2) * NB + b] * T[fcc];
    double den = W[(W_ZF + 3) * NB + b];
    if (MODE == MODE_RES) {
        // residual of den T_w = S, measured as the change the assignment
        // would make
        double r = S - den * T[fw];
        v.rf[(size_t)F_T * d.NF + fw] = (float)r;
        eT = fabs(r / den);
    } else {
        double nv = S / den;
        eT = fabs(nv - T[fw]);
        T[fw] = nv;
    }
}

template <int MODE>
__device__ __forceinline__ void T_coef(const Dev &d, const View<double> &v, int lb)
{
    int c = cell_of_block(d, lb, blockDim.x);
    bool act = c < d.N;
    double eT = 0.;
    if (act) {
        const int N = d.N;
        int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
        // the FP64 block; RES also keeps an FP32 copy for the corrections
        energy_coefs(d, c, i, j, k, [&](int slot, double x) {
            d.TF[slot * N + c] = x;
            if (MODE == MODE_RES) v.TF32[slot * N + c] = (float)x;
        });
        if (k == 0 || k == d.NZ - 1) zero_flux_wall<MODE>(d, v, i, j, k, eT);
    }
    err_commit(d, F_T, eT);
}

__global__ void __launch_bounds__(256) k_T_coef(Dev d)
{
    if (d.ctl->done) return;
    T_coef<MODE_GS>(d, view64(d), xcd_block());
}

// the adiabatic walls alone (the energy coefficients are then built inside
// the fused energy sweep)
__global__ void __launch_bounds__(256) k_zero_flux(Dev d)
{
    if (d.ctl->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int plane = d.NT * d.NR;
    double eT = 0.;
    if (t < 2 * plane) {
        const int q = t % plane, i = iof(d, q % d.NT), j = q / d.NT, k = t < plane ? 0 : d.NZ - 1;
        zero_flux_wall<MODE_GS>(d, view64(d), i, j, k, eT);
    }
    err_commit(d, F_T, eT);
}

// the energy coefficients alone at the current state, whatever the loop
// control says (the adjoint takes them at a converged state)
__global__ void __launch_bounds__(256) k_energy_coefs(Dev d)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.N) return;
    const int N = d.N;
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    energy_coefs(d, c, i, j, k, [&](int slot, double x) { d.TF[slot * N + c] = x; });
}

__global__ void __launch_bounds__(256) k_res_T_coef(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    T_coef<MODE_RES>(d, v, xcd_block());
}

// ---------------------------------------------------------------------
// sources (computed from the state at the start of each stage)
// ---------------------------------------------------------------------
// update_velocity_sources (nse.f90:199-209): calculate_gforces, calculate_R_b,
// calculate_R_D2. The lagged-convection term calculate_R_c_new
// (nse.f90:1393-1434) vanishes identically: meshes_u and uc are the same
// expression of the same time-centred velocity (nim.f90:169-171 and
// nse.f90:1422-1428), so it is not evaluated.
// One thread per cell for all three components, so that the cell's index
// arithmetic and its metrics (G, GO, 1/V, the p faces) are loaded once: the
// kernel is bound by the number of memory requests, not by HBM bytes (one
// wave per component, sharing those loads through the cache, took 15% longer).
// For corrections (v.old null) the previous-level part of the buoyancy is
// dropped: every other term is linear in the unknowns.
template <typename T>
__device__ __forceinline__ void mom_src_cell(const Dev &d, const View<T> &v, int lb)
{
    int c = cell_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    const int N = d.N;
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    int f[6];
    cell_faces(d, i, j, k, f);
    const T *P = v.fc + (size_t)F_P * d.NF;
    T pf[6];
#pragma unroll
    for (int p = 0; p < 6; p++) pf[p] = P[f[p]];
    const T invV = v.gc[G_INVV * N + c];
    T G[3], GO[6];
    cross_metrics(d, v.gc, c, G, GO);
    const T bw = -(T(d.g_beta) * T(0.5) * (v.sc[3 * N + c] + (v.old ? v.old[3 * N + c] : T(0))));
#pragma unroll
    for (int q = 0; q < 3; q++) {
        T RB = (q == 2 ? bw : T(0)) + T(d.rinv) * dphi(d, v.An, pf, f, q, invV);
        T RD = cross_diffusion(d, G, GO, v.fc + (size_t)q * d.NF, i, j, k, T(d.Rinv));
        v.src[q * N + c] = RD + RB;
    }
}

__global__ void __launch_bounds__(256) k_mom_src(Dev d)
{
    if (d.ctl->done) return;
    mom_src_cell(d, view64(d), xcd_block());
}

__global__ void __launch_bounds__(256) k_d_mom_src(Dev d, View<float> v)
{
    if (d.ctl->done) return;
    mom_src_cell(d, v, xcd_block());
}

// mom_src_cell for two cells of one half-row per thread (CellPair): the
// same arithmetic per cell, with one 16-byte load per pair of values
template <typename T>
__device__ __forceinline__ void mom_src_pair(const Dev &d, const View<T> &v, int lb)
{
    const int c = pair_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    const int N = d.N;
    const CellPair p = cell_pair(d, c);
    T pa[6], pb[6];
    faces2(d, p, v.fc + (size_t)F_P * d.NF, pa, pb);
    T invA, invB;
    ld2(v.gc + (size_t)G_INVV * N + c, invA, invB);
    T GA[3], GOA[6], GB[3], GOB[6];
    cross_metrics2(d, v.gc, c, GA, GOA, GB, GOB);
    T sa, sb, oa = T(0), ob = T(0);
    ld2(v.sc + 3 * (size_t)N + c, sa, sb);
    if (v.old) ld2(v.old + 3 * (size_t)N + c, oa, ob);
    const T bwa = -(T(d.g_beta) * T(0.5) * (sa + oa)), bwb = -(T(d.g_beta) * T(0.5) * (sb + ob));
#pragma unroll
    for (int q = 0; q < 3; q++) {
        T Aa[6], Ab[6];
        faces2(d, p, v.An + (size_t)q * d.NF, Aa, Ab);
        const T RBa = (q == 2 ? bwa : T(0)) + T(d.rinv) * dphi_vals(pa, Aa, invA);
        const T RBb = (q == 2 ? bwb : T(0)) + T(d.rinv) * dphi_vals(pb, Ab, invB);
        T RDa, RDb;
        cross_diffusion2(d, p, v.fc + (size_t)q * d.NF, GA, GOA, GB, GOB, T(d.Rinv), RDa, RDb);
        st2(v.src + (size_t)q * N + c, RDa + RBa, RDb + RBb);
    }
}

__global__ void __launch_bounds__(256) k_mom_src2(Dev d)
{
    if (d.ctl->done) return;
    mom_src_pair(d, view64(d), xcd_block());
}


// update_pressure_sources (nse.f90:228-235): calculate_R_b_p
// (nse.f90:2362-2407; faces_bx = faces_by = 0, faces_bz = -g_beta*faces_T)
// and calculate_R_D_p2; stores R_D_p + R_b_p, the combination the pressure
// sweep and wall update use, and (FP64) the dilatation. R_b_p is quadratic
// in the velocity gradients: the residual pass keeps the gradients at x0
// and the corrections use the exact expansion R_b_p(x0 + dx) - R_b_p(x0)
// (e.g. (a0 + da)^2 - a0^2 = 2 a0 da + da^2), which is small in dx.
template <int MODE, typename T>
__device__ __forceinline__ T ppe_rbp(const Dev &d, const View<T> &v, int c, int i, int j, int k)
{
    const int N = d.N;
    int f[6];
    cell_faces(d, i, j, k, f);
    T uf[6], vf[6], wf[6], bzf[6];
    const T *U = v.fc, *V = v.fc + d.NF, *W = v.fc + 2 * (size_t)d.NF, *Tf = v.fc + 4 * (size_t)d.NF;
    for (int p = 0; p < 6; p++) {
        uf[p] = U[f[p]];
        vf[p] = V[f[p]];
        wf[p] = W[f[p]];
        bzf[p] = -(T(d.g_beta) * Tf[f[p]]);
    }
    T invV = v.gc[G_INVV * N + c];
    T ux = dphi(d, v.An, uf, f, 0, invV), uy = dphi(d, v.An, uf, f, 1, invV), uz = dphi(d, v.An, uf, f, 2, invV);
    T vx = dphi(d, v.An, vf, f, 0, invV), vy = dphi(d, v.An, vf, f, 1, invV), vz = dphi(d, v.An, vf, f, 2, invV);
    T wx = dphi(d, v.An, wf, f, 0, invV), wy = dphi(d, v.An, wf, f, 1, invV), wz = dphi(d, v.An, wf, f, 2, invV);
    T bzz = dphi(d, v.An, bzf, f, 2, invV);
    T Dila = ux + vy + wz;
    T Rbp;
    if (MODE == MODE_DELTA) {
        const float *g0 = v.grad;
        T ux0 = g0[0 * N + c], uy0 = g0[1 * N + c], uz0 = g0[2 * N + c];
        T vx0 = g0[3 * N + c], vy0 = g0[4 * N + c], vz0 = g0[5 * N + c];
        T wx0 = g0[6 * N + c], wy0 = g0[7 * N + c], wz0 = g0[8 * N + c];
        T sq = (T(2) * ux0 + ux) * ux + (T(2) * vy0 + vy) * vy + (T(2) * wz0 + wz) * wz;
        T cr = uy0 * vx + uy * vx0 + uy * vx + vz0 * wy + vz * wy0 + vz * wy + wx0 * uz + wx * uz0 + wx * uz;
        Rbp = -sq - T(2) * cr - bzz + Dila / (T(2) * T(d.DT));
    } else {
        if (MODE == MODE_GS) v.sc[S_DILA * N + c] = Dila;
        if (MODE == MODE_RES) {
            float *g0 = v.grad;
            g0[0 * N + c] = (float)ux; g0[1 * N + c] = (float)uy; g0[2 * N + c] = (float)uz;
            g0[3 * N + c] = (float)vx; g0[4 * N + c] = (float)vy; g0[5 * N + c] = (float)vz;
            g0[6 * N + c] = (float)wx; g0[7 * N + c] = (float)wy; g0[8 * N + c] = (float)wz;
        }
        Rbp = -(ux * ux + vy * vy + wz * wz) - T(2) * (uy * vx + vz * wy + wx * uz) - bzz + Dila / (T(2) * T(d.DT));
    }
    return Rbp;
}

template <int MODE, typename T>
__device__ __forceinline__ void ppe_src(const Dev &d, const View<T> &v, int lb)
{
    int c = cell_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    const int N = d.N;
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    T RDp = cross_diffusion(d, v.gc, v.fc + (size_t)F_P * d.NF, i, j, k, c, T(d.rinv));
    const T Rbp = ppe_rbp<MODE>(d, v, c, i, j, k);
    v.src[R_P * N + c] = RDp + Rbp;
}

__global__ void __launch_bounds__(256) k_ppe_src(Dev d)
{
    if (d.ctl->done) return;
    ppe_src<MODE_GS>(d, view64(d), xcd_block());
}

__global__ void __launch_bounds__(256) k_res_ppe_src(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    ppe_src<MODE_RES>(d, v, xcd_block());
}

__global__ void __launch_bounds__(256) k_d_ppe_src(Dev d, View<float> v)
{
    if (d.ctl->done) return;
    ppe_src<MODE_DELTA>(d, v, xcd_block());
}

// ppe_rbp (GS and RES) of one cell from its loaded face values (u, v, w,
// T) and area components (A[comp][face]); the same operations as ppe_rbp
template <int MODE, typename T>
__device__ __forceinline__ T rbp_vals(const Dev &d, const View<T> &v, int c, const T uf[6], const T vf[6],
                                      const T wf[6], const T tf[6], const T A[3][6], T invV)
{
    static_assert(MODE != MODE_DELTA, "corrections use ppe_rbp");
    T bzf[6];
#pragma unroll
    for (int q = 0; q < 6; q++) bzf[q] = -(T(d.g_beta) * tf[q]);
    T ux = dphi_vals(uf, A[0], invV), uy = dphi_vals(uf, A[1], invV), uz = dphi_vals(uf, A[2], invV);
    T vx = dphi_vals(vf, A[0], invV), vy = dphi_vals(vf, A[1], invV), vz = dphi_vals(vf, A[2], invV);
    T wx = dphi_vals(wf, A[0], invV), wy = dphi_vals(wf, A[1], invV), wz = dphi_vals(wf, A[2], invV);
    T bzz = dphi_vals(bzf, A[2], invV);
    T Dila = ux + vy + wz;
    const int N = d.N;
    if (MODE == MODE_GS) v.sc[S_DILA * N + c] = Dila;
    if (MODE == MODE_RES) {
        float *gr = v.grad;
        gr[0 * N + c] = (float)ux; gr[1 * N + c] = (float)uy; gr[2 * N + c] = (float)uz;
        gr[3 * N + c] = (float)vx; gr[4 * N + c] = (float)vy; gr[5 * N + c] = (float)vz;
        gr[6 * N + c] = (float)wx; gr[7 * N + c] = (float)wy; gr[8 * N + c] = (float)wz;
    }
    return -(ux * ux + vy * vy + wz * wz) - T(2) * (uy * vx + vz * wy + wx * uz) - bzz + Dila / (T(2) * T(d.DT));
}

// ppe_src (GS and RES) for two cells of one half-row per thread (CellPair)
template <int MODE, typename T>
__device__ __forceinline__ void ppe_src_pair(const Dev &d, const View<T> &v, int lb)
{
    const int c = pair_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    const int N = d.N;
    const CellPair p = cell_pair(d, c);
    T GA[3], GOA[6], GB[3], GOB[6];
    cross_metrics2(d, v.gc, c, GA, GOA, GB, GOB);
    T RDa, RDb;
    cross_diffusion2(d, p, v.fc + (size_t)F_P * d.NF, GA, GOA, GB, GOB, T(d.rinv), RDa, RDb);
    T ua[6], ub[6], va[6], vb[6], wa[6], wb[6], ta[6], tb[6], Aa[3][6], Ab[3][6];
    faces2(d, p, v.fc, ua, ub);
    faces2(d, p, v.fc + (size_t)d.NF, va, vb);
    faces2(d, p, v.fc + 2 * (size_t)d.NF, wa, wb);
    faces2(d, p, v.fc + 4 * (size_t)d.NF, ta, tb);
#pragma unroll
    for (int q = 0; q < 3; q++) faces2(d, p, v.An + (size_t)q * d.NF, Aa[q], Ab[q]);
    T invA, invB;
    ld2(v.gc + (size_t)G_INVV * N + c, invA, invB);
    const T Ra = rbp_vals<MODE>(d, v, c, ua, va, wa, ta, Aa, invA);
    const T Rb = rbp_vals<MODE>(d, v, c + 1, ub, vb, wb, tb, Ab, invB);
    st2(v.src + (size_t)R_P * N + c, RDa + Ra, RDb + Rb);
}

__global__ void __launch_bounds__(256) k_ppe_src2(Dev d)
{
    if (d.ctl->done) return;
    ppe_src_pair<MODE_GS>(d, view64(d), xcd_block());
}

__global__ void __launch_bounds__(256) k_res_ppe_src2(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    ppe_src_pair<MODE_RES>(d, v, xcd_block());
}


// update_temperature_sources (nse.f90:249-256): R_b_T = q = 0, R_D_T
template <typename T>
__device__ __forceinline__ void T_src(const Dev &d, const View<T> &v, int lb)
{
    int c = cell_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    v.src[R_T * d.N + c] = cross_diffusion(d, v.gc, v.fc + (size_t)F_T * d.NF, i, j, k, c, T(d.alpha)) + T(0);
}

__global__ void __launch_bounds__(256) k_T_src(Dev d)
{
    if (d.ctl->done) return;
    T_src(d, view64(d), xcd_block());
}

// T_src for two cells of one half-row per thread (CellPair)
template <typename T>
__device__ __forceinline__ void T_src_pair(const Dev &d, const View<T> &v, int lb)
{
    const int c = pair_of_block(d, lb, blockDim.x);
    if (c >= d.N) return;
    const CellPair p = cell_pair(d, c);
    T GA[3], GOA[6], GB[3], GOB[6];
    cross_metrics2(d, v.gc, c, GA, GOA, GB, GOB);
    T ra, rb;
    cross_diffusion2(d, p, v.fc + (size_t)F_T * d.NF, GA, GOA, GB, GOB, T(d.alpha), ra, rb);
    st2(v.src + (size_t)R_T * d.N + c, ra + T(0), rb + T(0));
}

__global__ void __launch_bounds__(256) k_T_src2(Dev d)
{
    if (d.ctl->done) return;
    T_src_pair(d, view64(d), xcd_block());
}


// the same from an LDS-staged tile of 256 cells (NT a multiple of TI)
template <int TI>
__global__ void __launch_bounds__(256) k_T_src_tile(Dev d)
{
    if (d.ctl->done) return;
    __shared__ FaceTile<TI> s;
    int i0, j0, k;
    tile_of_block<TI>(d, xcd_block(), i0, j0, k);
    const double *F = d.fc + (size_t)F_T * d.NF;
    load_tile<TI>(d, F, s, i0, j0, k);
    __syncthreads();
    const int li = threadIdx.x % TI, jr = threadIdx.x / TI, i = i0 + li, j = j0 + jr;
    if (j >= d.NR) return;
    const int c = cidx(d, i, j, k);
    d.src[R_T * d.N + c] = cross_diffusion_tile<TI>(d, s, F, i, j, k, c, li, jr, j0, d.alpha) + 0.;
}
template __global__ void k_T_src_tile<64>(Dev);
template __global__ void k_T_src_tile<32>(Dev);

// the momentum sources (mom_src_cell) with the cross-diffusion of u, v, w
// from LDS-staged tiles of one plane (three tiles, 46 KB at 64 x 4 cells)
template <int TI>
__global__ void __launch_bounds__(256) k_mom_src_tile(Dev d)
{
    if (d.ctl->done) return;
    __shared__ FaceTile<TI> s[3];
    int i0, j0, k;
    tile_of_block<TI>(d, xcd_block(), i0, j0, k);
    for (int q = 0; q < 3; q++) load_tile<TI>(d, d.fc + (size_t)q * d.NF, s[q], i0, j0, k);
    __syncthreads();
    const int li = threadIdx.x % TI, jr = threadIdx.x / TI, i = i0 + li, j = j0 + jr;
    if (j >= d.NR) return;
    const int c = cidx(d, i, j, k), N = d.N;
    int f[6];
    cell_faces(d, i, j, k, f);
    const double *P = d.fc + (size_t)F_P * d.NF;
    double pf[6];
#pragma unroll
    for (int p = 0; p < 6; p++) pf[p] = P[f[p]];
    const double invV = d.gc[G_INVV * N + c];
    const double bw = -(d.g_beta * 0.5 * (d.sc[3 * N + c] + d.sc[(size_t)S_OLD * N + 3 * N + c]));
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double RB = (q == 2 ? bw : 0.) + d.rinv * dphi(d, d.An, pf, f, q, invV);
        const double RD = cross_diffusion_tile<TI>(d, s[q], d.fc + (size_t)q * d.NF, i, j, k, c, li, jr, j0, d.Rinv);
        d.src[q * N + c] = RD + RB;
    }
}
template __global__ void k_mom_src_tile<64>(Dev);
template __global__ void k_mom_src_tile<32>(Dev);

// the pressure sources (ppe_src) with the cross-diffusion of p from an
// LDS-staged tile
template <int TI>
__global__ void __launch_bounds__(256) k_ppe_src_tile(Dev d)
{
    if (d.ctl->done) return;
    __shared__ FaceTile<TI> s;
    int i0, j0, k;
    tile_of_block<TI>(d, xcd_block(), i0, j0, k);
    const double *F = d.fc + (size_t)F_P * d.NF;
    load_tile<TI>(d, F, s, i0, j0, k);
    __syncthreads();
    const int li = threadIdx.x % TI, jr = threadIdx.x / TI, i = i0 + li, j = j0 + jr;
    if (j >= d.NR) return;
    const int c = cidx(d, i, j, k);
    const double Rbp = ppe_rbp<MODE_GS>(d, view64(d), c, i, j, k);
    const double RDp = cross_diffusion_tile<TI>(d, s, F, i, j, k, c, li, jr, j0, d.rinv);
    d.src[R_P * d.N + c] = RDp + Rbp;
}
template __global__ void k_ppe_src_tile<64>(Dev);
template __global__ void k_ppe_src_tile<32>(Dev);

__global__ void __launch_bounds__(256) k_d_T_src(Dev d, View<float> v)
{
    if (d.ctl->done) return;
    T_src(d, v, xcd_block());
}

// ---------------------------------------------------------------------
// relaxation sweeps
// ---------------------------------------------------------------------
// one transported field of one cell: the cell average, then owned faces
// 2, 4, 6 (integrate_velocity_field_formulation2_i nse.f90:831-1006 and
// integrate_temperature_field_formulation2_i nse.f90:1009-1089).
// Gather, compute, scatter: every input is loaded before anything is
// stored, and the unit's own updated values (cell, face 2, face 4) are
// carried in registers to the later updates that read them, as the
// reference's in-place order does. No other cell of the colour writes any
// location read here, so loading early changes no value.
// ff is the face field (F_U..F_T), cs the cell slot (0..3). The
// coefficients come from the block coef ([NCF][N]) or, with REG, from the
// caller's registers creg[NCF]; the metric differences from the block or,
// when given, from the caller's registers greg[6].
template <int MODE, typename T, bool REG = false, bool NTL = false, bool NTG = NTL && !REG>
__device__ void transport_unit(const Dev &d, const View<T> &vw, int ff, int cs, const T *__restrict__ coef, T B, T w,
                               int i, int j, int k, int c, double &ecell, double &eface, const T *creg = nullptr,
                               const T *greg = nullptr)
{
    const int N = d.N;
    T *__restrict__ Fc = vw.fc + (size_t)ff * d.NF;
    T *__restrict__ C = vw.sc + (size_t)cs * N;
    const T *__restrict__ CO = MODE == MODE_DELTA ? nullptr : vw.old + (size_t)cs * N;
    const bool has2 = j < d.NR - 1, has6 = k < d.NZ - 1;
    // neighbour rows clamped at the walls: the loads stay in bounds and the
    // results are discarded there
    const int jn = has2 ? j + 1 : j, kn = has6 ? k + 1 : k, ip = ip1(d, i);
    const T *gdf = vw.gc + (size_t)G_DIFF * N;
    int f[6];
    cell_faces(d, i, j, k, f);
    // ---- gather
    const T old = C[c], oc = MODE == MODE_DELTA ? T(0) : CO[c];
    // The diagonals F4(1), F1(1), F2(1), F3(1) are sums of the other
    // coefficients (nim.f90:292-329, 465-502). They are summed again here in
    // the order the coefficient kernels use, which saves four loads per cell:
    // in FP64 that gives the stored values bit for bit; the FP32 correction
    // passes sum the FP32 copies and may differ from the rounded FP64
    // diagonal in the last bit, a rounding of the correction's smoother only
    // (the FP64 residual fixes what the corrections converge to).
    constexpr bool sum_diag = true;
    constexpr int m0 = sum_diag ? 1 : 0;
    // NTL: non-temporal loads of the coefficients, residuals and (NTG) the
    // metric differences. In the FP32 momentum sweep the three waves of a
    // block share these lines: the sweep itself slows by 6% but the source
    // kernels after it gain more (one FP32 pass 5% faster).
    T a[8];
#pragma unroll
    for (int m = REG ? 0 : m0; m < 8; m++) a[m] = REG ? creg[C_F4 + m] : ld_pass<NTL>(&coef[(C_F4 + m) * N + c]);
    T F0 = Fc[f[0]], F1 = Fc[f[1]], F2 = Fc[f[2]], F3 = Fc[f[3]], F4 = Fc[f[4]], F5 = Fc[f[5]];
    T b[5], e[5], h[5];
#pragma unroll
    for (int m = REG ? 0 : m0; m < 5; m++) {
        b[m] = REG ? creg[C_F1 + m] : ld_pass<NTL>(&coef[(C_F1 + m) * N + c]);
        e[m] = REG ? creg[C_F2 + m] : ld_pass<NTL>(&coef[(C_F2 + m) * N + c]);
        h[m] = REG ? creg[C_F3 + m] : ld_pass<NTL>(&coef[(C_F3 + m) * N + c]);
    }
    if (sum_diag && !REG) {
        a[0] = a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7];
        b[0] = b[1] + b[2] + T(2) * b[3] + T(2) * b[4];
        e[0] = e[1] + e[2] + T(2) * e[3] + T(2) * e[4];
        h[0] = h[1] + h[2] + T(2) * h[3] + T(2) * h[4];
    }
    T gd[6];
#pragma unroll
    for (int m = 0; m < 6; m++) gd[m] = greg ? greg[m] : ld_pass<NTG>(&gdf[m * N + c]);
    const T gd0 = gd[0], gd1 = gd[1], gd2 = gd[2], gd3 = gd[3], gd4 = gd[4], gd5 = gd[5];
    // face 2 partner: cell (i, j+1, k)
    const int cr = cidx(d, i, jn, k);
    const T r2 = Fc[fidx(d, 2, i, jn, k)], r3 = Fc[fidx(d, 3, i, jn, k)], r4 = Fc[fidx(d, 4, i, jn, k)];
    const T r5 = Fc[fidx(d, 5, i, jn, k)], r6 = Fc[fidx(d, 6, i, jn, k)];
    const T sr = MODE == MODE_DELTA ? C[cr] : C[cr] + CO[cr];
    // face 4 partner: cell (i+1, j, k), periodic
    const int ct = cidx(d, ip, j, k);
    const T t1 = Fc[fidx(d, 1, ip, j, k)], t2 = Fc[fidx(d, 2, ip, j, k)], t4 = Fc[fidx(d, 4, ip, j, k)];
    const T t5 = Fc[fidx(d, 5, ip, j, k)], t6 = Fc[fidx(d, 6, ip, j, k)];
    const T st = MODE == MODE_DELTA ? C[ct] : C[ct] + CO[ct];
    // face 6 partner: cell (i, j, k+1)
    const int cz = cidx(d, i, j, kn);
    const T z1 = Fc[fidx(d, 1, i, j, kn)], z2 = Fc[fidx(d, 2, i, j, kn)], z3 = Fc[fidx(d, 3, i, j, kn)];
    const T z4 = Fc[fidx(d, 4, i, j, kn)], z6 = Fc[fidx(d, 6, i, j, kn)];
    const T sz = MODE == MODE_DELTA ? C[cz] : C[cz] + CO[cz];
    float *rF = MODE == MODE_GS ? nullptr : vw.rf + (size_t)ff * d.NF;
    float *rC = MODE == MODE_GS ? nullptr : vw.rc + (size_t)cs * N;
    T q0 = T(0), q2 = T(0), q4 = T(0), q6 = T(0);
    if (MODE == MODE_DELTA) {
        q0 = ld_pass<NTL>(&rC[c]);
        q2 = ld_pass<NTL>(&rF[f[1]]);
        q4 = ld_pass<NTL>(&rF[f[3]]);
        q6 = ld_pass<NTL>(&rF[f[5]]);
    }
    // ---- cell average
    const T S0 = a[1] * oc + a[2] * F0 + a[3] * F1 + a[4] * F2 + a[5] * F3 + a[6] * F4 + a[7] * F5 + B;
    T nv = old;
    if (MODE == MODE_RES) {
        T r = S0 - a[0] * old;
        rC[c] = (float)r;
        ecell = err_max(ecell, ((double)(w * r / a[0])));
    } else {
        nv = (1 - w) * old + (w / a[0]) * (MODE == MODE_DELTA ? S0 + q0 : S0);
        ecell = err_max(ecell, ((double)(nv - old)));
    }
    const T ccur = nv + oc;
    // ---- face 2
    if (has2) {
        T B3 = gd0 * (F3 - F2 + r4 - r3) / T(4);
        T B4 = gd1 * (F5 - F4 + r6 - r5) / T(4);
        T S = b[1] * F0 + b[2] * r2 + b[3] * ccur + b[4] * sr - B3 - B4;
        if (MODE == MODE_RES) {
            T r = S - b[0] * F1;
            rF[f[1]] = (float)r;
            eface = err_max(eface, ((double)(w * r / b[0])));
        } else {
            T v = (1 - w) * F1 + (w / b[0]) * (MODE == MODE_DELTA ? S + q2 : S);
            eface = err_max(eface, ((double)(v - F1)));
            F1 = v;
        }
    }
    // ---- face 4
    {
        T B3 = gd2 * (F1 - F0 + t2 - t1) / T(4);
        T B4 = gd3 * (F5 - F4 + t6 - t5) / T(4);
        T S = e[1] * F2 + e[2] * t4 + e[3] * ccur + e[4] * st - B3 - B4;
        if (MODE == MODE_RES) {
            T r = S - e[0] * F3;
            rF[f[3]] = (float)r;
            eface = err_max(eface, ((double)(w * r / e[0])));
        } else {
            T v = (1 - w) * F3 + (w / e[0]) * (MODE == MODE_DELTA ? S + q4 : S);
            eface = err_max(eface, ((double)(v - F3)));
            F3 = v;
        }
    }
    // ---- face 6
    if (has6) {
        T B3 = gd4 * (F1 - F0 + z2 - z1) / T(4);
        T B4 = gd5 * (F3 - F2 + z4 - z3) / T(4);
        T S = h[1] * F4 + h[2] * z6 + h[3] * ccur + h[4] * sz - B3 - B4;
        if (MODE == MODE_RES) {
            T r = S - h[0] * F5;
            rF[f[5]] = (float)r;
            eface = err_max(eface, ((double)(w * r / h[0])));
        } else {
            T v = (1 - w) * F5 + (w / h[0]) * (MODE == MODE_DELTA ? S + q6 : S);
            eface = err_max(eface, ((double)(v - F5)));
            F5 = v;
        }
    }
    // ---- scatter
    if (MODE != MODE_RES) {
        C[c] = nv;
        if (has2) Fc[f[1]] = F1;
        Fc[f[3]] = F3;
        if (has6) Fc[f[5]] = F5;
    }
}

// cell of this thread: an item of a sweep schedule, or (RES) every cell
template <int MODE>
__device__ __forceinline__ bool sweep_cell(const Dev &d, const Sched &s, int lb, int &i, int &j, int &k)
{
    if (MODE != MODE_RES) return sched_cell(d, s, lb, i, j, k);
    int tid = lb * blockDim.x + threadIdx.x;
    if (tid >= d.N) return false;
    i = iof(d, tid % d.NT);
    j = (tid / d.NT) % d.NR;
    k = tid / (d.NT * d.NR);
    return true;
}

// residual pass and FP32 corrections of the mixed loop: threadIdx.y selects
// the field (u, v, w are independent within the sweep)
template <int MODE, typename T, bool NTL = false>
__device__ __forceinline__ void mom_sweep(const Dev &d, const View<T> &v, const Sched &sc, int lb)
{
    int i, j, k;
    const int q = threadIdx.y;
    bool act = sweep_cell<MODE>(d, sc, lb, i, j, k);
    double e = 0.;
    if (act) {
        const int N = d.N;
        int c = cidx(d, i, j, k);
        const T w = T(q == 0 ? d.w_u : q == 1 ? d.w_v : d.w_w);
        double ec = 0., ef = 0.;
        transport_unit<MODE, T, false, NTL>(d, v, q, q, v.MF, v.src[q * N + c], w, i, j, k, c, ec, ef);
        e = fmax(ec, ef);
    }
    if (MODE != MODE_DELTA) err_commit(d, q, e);
}

// The same with one thread per cell for the three components: the cell's
// momentum coefficients M_F (shared by u, v and w; in FP64 with the
// diagonals re-summed) and its metric differences are loaded once and handed
// to each unit in registers (the units' stores would otherwise make each
// reload them).
template <int MODE, typename T, bool NTL = false>
__device__ __forceinline__ void mom_sweep_cell(const Dev &d, const View<T> &v, const Sched &sc, int lb)
{
    int i, j, k;
    bool act = sweep_cell<MODE>(d, sc, lb, i, j, k);
    double e[3] = {0., 0., 0.};
    if (act) {
        const int N = d.N;
        const int c = cidx(d, i, j, k);
        const T *__restrict__ coef = v.MF;
        T creg[NCF];
#pragma unroll
        for (int m = 0; m < NCF; m++) creg[m] = ld_pass<NTL>(&coef[m * N + c]);
        if (sizeof(T) == sizeof(double)) {
            // the FP64 sweep re-sums the diagonals as transport_unit does
            T *a = creg + C_F4;
            a[0] = a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7];
#pragma unroll
            for (int b0 = C_F1; b0 <= C_F3; b0 += 5)
                creg[b0] = creg[b0 + 1] + creg[b0 + 2] + T(2) * creg[b0 + 3] + T(2) * creg[b0 + 4];
        }
        T greg[6];
#pragma unroll
        for (int m = 0; m < 6; m++) greg[m] = v.gc[(G_DIFF + m) * N + c];
        const T w[3] = {T(d.w_u), T(d.w_v), T(d.w_w)};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double ec = 0., ef = 0.;
            transport_unit<MODE, T, true>(d, v, q, q, coef, v.src[q * N + c], w[q], i, j, k, c, ec, ef, creg, greg);
            e[q] = fmax(ec, ef);
        }
    }
    if (MODE != MODE_DELTA)
#pragma unroll
        for (int q = 0; q < 3; q++) err_commit(d, q, e[q]);
}

template <bool NTL>
__global__ void __launch_bounds__(256) k_mom_sweep(Dev d, Sched sc)
{
    if (d.ctl->done) return;
    mom_sweep_cell<MODE_GS, double, NTL>(d, view64(d), sc, xcd_block());
}
template __global__ void k_mom_sweep<false>(Dev, Sched);
template __global__ void k_mom_sweep<true>(Dev, Sched);

__global__ void __launch_bounds__(192) k_res_mom(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    mom_sweep<MODE_RES>(d, v, Sched{}, xcd_block());
}

// FP32 corrections: one wave per component (the FP32 unit is light on
// registers, and this runs 2% faster than one thread per cell)
template <bool NTL>
__global__ void __launch_bounds__(192) k_d_mom_sweep(Dev d, View<float> v, Sched sc)
{
    if (d.ctl->done) return;
    mom_sweep<MODE_DELTA, float, NTL>(d, v, sc, xcd_block());
}
template __global__ void k_d_mom_sweep<false>(Dev, View<float>, Sched);
template __global__ void k_d_mom_sweep<true>(Dev, View<float>, Sched);

template <int MODE, typename T, bool NTL = false>
__device__ __forceinline__ void T_sweep(const Dev &d, const View<T> &v, const T *coef, const Sched &sc, int lb)
{
    int i, j, k;
    bool act = sweep_cell<MODE>(d, sc, lb, i, j, k);
    double e = 0.;
    if (act) {
        const int N = d.N;
        int c = cidx(d, i, j, k);
        double ec = 0., ef = 0.;
        transport_unit<MODE, T, false, NTL>(d, v, F_T, 3, coef, v.src[R_T * N + c], T(d.w_T), i, j, k, c, ec, ef);
        e = fmax(ec, ef);
    }
    if (MODE != MODE_DELTA) err_commit(d, F_T, e);
}

__global__ void __launch_bounds__(256) k_T_sweep(Dev d, Sched sc)
{
    if (d.ctl->done) return;
    View<double> v = view64(d);
    T_sweep<MODE_GS>(d, v, v.TF, sc, xcd_block());
}

// the energy sweep with its coefficients built in registers from the
// velocity (T_coef's formulas), for the FP64 loop: the coefficients never
// go through memory
template <bool NTL>
__global__ void __launch_bounds__(256) k_T_sweep_fused(Dev d, Sched sc)
{
    if (d.ctl->done) return;
    int i, j, k;
    bool act = sched_cell(d, sc, xcd_block(), i, j, k);
    double e = 0.;
    if (act) {
        const int N = d.N, c = cidx(d, i, j, k);
        // slots without a + neighbour stay 0: the unit does not use them
        double F[NCF];
#pragma unroll
        for (int m = 0; m < NCF; m++) F[m] = 0.;
        energy_coefs(d, c, i, j, k, [&](int slot, double x) { F[slot] = x; });
        const View<double> v = view64(d);
        double ec = 0., ef = 0.;
        transport_unit<MODE_GS, double, true, NTL, NTL>(d, v, F_T, 3, nullptr, v.src[R_T * N + c], d.w_T, i, j, k, c, ec, ef, F);
        e = fmax(ec, ef);
    }
    err_commit(d, F_T, e);
}
template __global__ void k_T_sweep_fused<false>(Dev, Sched);
template __global__ void k_T_sweep_fused<true>(Dev, Sched);

__global__ void __launch_bounds__(256) k_res_T(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    T_sweep<MODE_RES>(d, v, v.TF, Sched{}, xcd_block());
}

template <bool NTL>
__global__ void __launch_bounds__(256) k_d_T_sweep(Dev d, View<float> v, Sched sc)
{
    if (d.ctl->done) return;
    T_sweep<MODE_DELTA, float, NTL>(d, v, v.TF, sc, xcd_block());
}
template __global__ void k_d_T_sweep<false>(Dev, View<float>, Sched);
template __global__ void k_d_T_sweep<true>(Dev, View<float>, Sched);

// apply_pressure_bc for the wall face at position pos of cell (i,j,k)
// (bc.f90:147-341). Every term is linear in the unknowns (p, u, v, w, T and
// the pressure source), so the correction form is the same formula on the
// corrections plus the residual.
template <int MODE, typename T>
__device__ void pressure_wall(const Dev &d, const View<T> &vw, int b, int pos, int i, int j, int k, int c, PErr &e)
{
    const int NB = d.NB;
    const T *W = vw.wall;
    T *P = vw.fc + (size_t)F_P * d.NF;
    const T *U = vw.fc, *V = vw.fc + d.NF, *Wv = vw.fc + 2 * (size_t)d.NF, *Tf = vw.fc + 4 * (size_t)d.NF;
    T n1 = W[(W_N + 0) * NB + b], n2 = W[(W_N + 1) * NB + b], n3 = W[(W_N + 2) * NB + b];
    int fw = fidx(d, pos, i, j, k);
    T b_n = T(0) * n1 + T(0) * n2 + (-(T(d.g_beta) * Tf[fw])) * n3;
    T d2 = T(0);
    for (int m = 0; m < 5; m++) {
        int q = d.wpts[m * NB + b];
        T vn = U[q] * n1 + V[q] * n2 + Wv[q] * n3;
        d2 = m == 0 ? -(W[(W_A + 0) * NB + b] * vn) : d2 - W[(W_A + m) * NB + b] * vn;
    }
    T gB2 = W[W_GB2 * NB + b], gB2c = W[W_GB2C * NB + b], gB3 = W[W_GB3 * NB + b];
    T B2, B3;
    if (pos <= 2) {
        // tangential directions theta (always central) and z
        B2 = tder(d, (const T *)P, pos, i, j, k, 0) * gB2c;
        B3 = tder(d, (const T *)P, pos, i, j, k, 2) * gB3;
    } else {
        // tangential directions r and theta (always central); the r
        // derivative takes a different metric factor in its central branch
        B2 = tder(d, (const T *)P, pos, i, j, k, 1) * (j == 0 || j == d.NR - 1 ? gB2 : gB2c);
        B3 = tder(d, (const T *)P, pos, i, j, k, 0) * gB3;
    }
    T B1 = W[(W_AP + 7) * NB + b] * vw.src[R_P * d.N + c];
    T s = T(0);
    bool first = true;
    for (int q = 1; q <= 6; q++) {
        if (q == pos) continue;
        T t = W[(W_AP + q - 1) * NB + b] * P[fidx(d, q, i, j, k)];
        s = first ? -t : s - t;
        first = false;
    }
    const T o = P[fw], ap = W[(W_AP + pos - 1) * NB + b], w2 = T(d.w_p2);
    const T S = s + B1 - B2 - B3 + (T(d.nu) * T(d.rho) * d2 - b_n * T(d.rho));
    float *rF = MODE == MODE_GS ? nullptr : vw.rf + (size_t)F_P * d.NF;
    if (MODE == MODE_RES) {
        T r = S - ap * o;
        rF[fw] = (float)r;
        e.add((double)(w2 * r / ap));
    } else {
        T v = (1 - w2) * o + w2 / ap * (MODE == MODE_DELTA ? S + rF[fw] : S);
        P[fw] = v;
        if (MODE == MODE_GS) e.add((double)(v - o));
    }
}

// integrate_pressure_field_i (nse.f90:1091-1289) for one cell, preceded by
// the cell's own wall faces in the reference's boundary-group order.
// After the walls, every input of the three interior face updates is
// gathered before the first store; the cell's own faces are carried in
// registers, so face 4 sees the new face 2 and face 6 the new faces 2 and 4.
template <int MODE, typename T, bool NTL = false>
__device__ __forceinline__ void ppe_sweep(const Dev &d, const View<T> &vw, const Sched &sc, int lb)
{
    int i, j, k;
    bool act = sweep_cell<MODE>(d, sc, lb, i, j, k);
    PErr e;
    if (act) {
        const int N = d.N, NT = d.NT, NR = d.NR, NZ = d.NZ;
        int c = cidx(d, i, j, k);
        if (k == 0) pressure_wall<MODE>(d, vw, i + NT * j, 5, i, j, k, c, e);
        if (k == NZ - 1) pressure_wall<MODE>(d, vw, NT * NR + i + NT * j, 6, i, j, k, c, e);
        if (j == 0) pressure_wall<MODE>(d, vw, 2 * NT * NR + i + NT * k, 1, i, j, k, c, e);
        if (j == NR - 1) pressure_wall<MODE>(d, vw, 2 * NT * NR + NT * NZ + i + NT * k, 2, i, j, k, c, e);
        T *__restrict__ P = vw.fc + (size_t)F_P * d.NF;
        const T *__restrict__ gdf = vw.gc + (size_t)G_DIFF * N;
        const T *__restrict__ R = vw.src + (size_t)R_P * N;
        const T w = T(d.w_p);
        const bool hx = j < NR - 1, hz = k < NZ - 1;
        const int jn = hx ? j + 1 : j, kn = hz ? k + 1 : k, ip = ip1(d, i);
        int f[6];
        cell_faces(d, i, j, k, f);
        // ---- gather
        T pf[6];
#pragma unroll
        for (int p = 0; p < 6; p++) pf[p] = P[f[p]];
        const T S1 = R[c];
        const T *__restrict__ Q = vw.AQ;
        T qx[NAQ], qy[NAQ], qz[NAQ];
#pragma unroll
        for (int m = 0; m < NAQ; m++) {
            qx[m] = ld_pass<NTL>(&Q[m * N + c]);
            qy[m] = ld_pass<NTL>(&Q[(NAQ + m) * N + c]);
            qz[m] = ld_pass<NTL>(&Q[(2 * NAQ + m) * N + c]);
        }
        // x: faces of cell (i, j+1, k)
        const T Sx = R[cidx(d, i, jn, k)];
        const T x2 = P[fidx(d, 2, i, jn, k)], x3 = P[fidx(d, 3, i, jn, k)], x4 = P[fidx(d, 4, i, jn, k)];
        const T x5 = P[fidx(d, 5, i, jn, k)], x6 = P[fidx(d, 6, i, jn, k)];
        const T Bx3 = ld_pass<NTL>(&gdf[0 * N + c]) * tder(d, (const T *)P, 2, i, j, k, 0);
        const T Bx4 = ld_pass<NTL>(&gdf[1 * N + c]) * tder(d, (const T *)P, 2, i, j, k, 2);
        // y: faces of cell (i+1, j, k)
        const T Sy = R[cidx(d, ip, j, k)];
        const T y1 = P[fidx(d, 1, ip, j, k)], y2 = P[fidx(d, 2, ip, j, k)], y4 = P[fidx(d, 4, ip, j, k)];
        const T y5 = P[fidx(d, 5, ip, j, k)], y6 = P[fidx(d, 6, ip, j, k)];
        const T By3 = ld_pass<NTL>(&gdf[2 * N + c]) * tder(d, (const T *)P, 4, i, j, k, 1);
        const T By4 = ld_pass<NTL>(&gdf[3 * N + c]) * tder(d, (const T *)P, 4, i, j, k, 2);
        // z: faces of cell (i, j, k+1)
        const T Sz = R[cidx(d, i, j, kn)];
        const T z1 = P[fidx(d, 1, i, j, kn)], z2 = P[fidx(d, 2, i, j, kn)], z3 = P[fidx(d, 3, i, j, kn)];
        const T z4 = P[fidx(d, 4, i, j, kn)], z6 = P[fidx(d, 6, i, j, kn)];
        const T Bz3 = ld_pass<NTL>(&gdf[4 * N + c]) * tder(d, (const T *)P, 6, i, j, k, 1);
        const T Bz4 = ld_pass<NTL>(&gdf[5 * N + c]) * tder(d, (const T *)P, 6, i, j, k, 0);
        float *rF = MODE == MODE_GS ? nullptr : vw.rf + (size_t)F_P * d.NF;
        T q2 = T(0), q4 = T(0), q6 = T(0);
        if (MODE == MODE_DELTA) {
            q2 = ld_pass<NTL>(&rF[f[1]]);
            q4 = ld_pass<NTL>(&rF[f[3]]);
            q6 = ld_pass<NTL>(&rF[f[5]]);
        }
        // one face equation: relax in place, or write its residual
        auto face = [&](T &pv, int slot, T diag, T S, T q) {
            if (MODE == MODE_RES) {
                T r = S - diag * pv;
                rF[slot] = (float)r;
                e.add((double)(w * r / diag));
            } else {
                T v = (1 - w) * pv + w / diag * (MODE == MODE_DELTA ? S + q : S);
                if (MODE == MODE_GS) e.add((double)(v - pv));
                pv = v;
            }
        };
        // ---- compute: AA_p?(m) of nim.f90:622-697 as the Q_ slots
        if (hx) {
            T B1 = S1 * qx[Q_S1], B2 = Sx * qx[Q_S2];
            T S = -qx[Q_SELF] * pf[0] - qx[Q_FAR] * x2 - qx[Q_TA0] * pf[2] - qx[Q_TA0] * pf[3] - qx[Q_TB0] * x3
                  - qx[Q_TB0] * x4 - qx[Q_TA1] * pf[4] - qx[Q_TA1] * pf[5] - qx[Q_TB1] * x5 - qx[Q_TB1] * x6 + B1 + B2
                  + Bx3 + Bx4;
            face(pf[1], f[1], qx[Q_DIAG], S, q2);
        }
        {
            T B1 = S1 * qy[Q_S1], B2 = Sy * qy[Q_S2];
            T S = -qy[Q_TA0] * pf[0] - qy[Q_TA0] * pf[1] - qy[Q_TB0] * y1 - qy[Q_TB0] * y2 - qy[Q_SELF] * pf[2]
                  - qy[Q_FAR] * y4 - qy[Q_TA1] * pf[4] - qy[Q_TA1] * pf[5] - qy[Q_TB1] * y5 - qy[Q_TB1] * y6 + B1 + B2
                  + By3 + By4;
            face(pf[3], f[3], qy[Q_DIAG], S, q4);
        }
        if (hz) {
            T B1 = S1 * qz[Q_S1], B2 = Sz * qz[Q_S2];
            T S = -qz[Q_TA0] * pf[0] - qz[Q_TA0] * pf[1] - qz[Q_TB0] * z1 - qz[Q_TB0] * z2 - qz[Q_TA1] * pf[2]
                  - qz[Q_TA1] * pf[3] - qz[Q_TB1] * z3 - qz[Q_TB1] * z4 - qz[Q_SELF] * pf[4] - qz[Q_FAR] * z6 + B1 + B2
                  + Bz3 + Bz4;
            face(pf[5], f[5], qz[Q_DIAG], S, q6);
        }
        // ---- scatter
        if (MODE != MODE_RES) {
            if (hx) P[f[1]] = pf[1];
            P[f[3]] = pf[3];
            if (hz) P[f[5]] = pf[5];
        }
    }
    if (MODE != MODE_DELTA) err_commit_p(d, e);
}

template <bool NTL>
__global__ void __launch_bounds__(256) k_ppe_sweep(Dev d, Sched sc)
{
    if (d.ctl->done) return;
    ppe_sweep<MODE_GS, double, NTL>(d, view64(d), sc, xcd_block());
}
template __global__ void k_ppe_sweep<false>(Dev, Sched);
template __global__ void k_ppe_sweep<true>(Dev, Sched);

__global__ void __launch_bounds__(256) k_res_ppe(Dev d, View<double> v)
{
    if (d.ctl->done) return;
    ppe_sweep<MODE_RES>(d, v, Sched{}, xcd_block());
}

template <bool NTL>
__global__ void __launch_bounds__(256) k_d_ppe_sweep(Dev d, View<float> v, Sched sc)
{
    if (d.ctl->done) return;
    ppe_sweep<MODE_DELTA, float, NTL>(d, v, sc, xcd_block());
}
template __global__ void k_d_ppe_sweep<false>(Dev, View<float>, Sched);
template __global__ void k_d_ppe_sweep<true>(Dev, View<float>, Sched);

// ---------------------------------------------------------------------
// mixed precision: FP32 corrections between FP64 residuals
// ---------------------------------------------------------------------
// apply_zero_flux (bc.f90:345-390) on the corrections: the FP64 residual r
// of den T_w = S(T) closes it, dT_w = (S(dT) + r) / den
__device__ __forceinline__ void d_zero_flux(const Dev &d, const View<float> &v, int lb)
{
    int b = lb * blockDim.x + threadIdx.x;
    const int NT = d.NT, NR = d.NR, NZ = d.NZ, NB = d.NB;
    if (b >= 2 * NT * NR) return;
    float *T = v.fc + (size_t)F_T * d.NF;
    const float *rT = v.rf + (size_t)F_T * d.NF;
    int fw, fa, fb, fcc;
    if (b < NT * NR) {
        int i = b % NT, j = b / NT;
        fw = fidx(d, 5, i, j, 0); fa = fidx(d, 6, i, j, 0); fb = fidx(d, 6, i, j, 1); fcc = fidx(d, 6, i, j, 2);
    } else {
        int i = (b - NT * NR) % NT, j = (b - NT * NR) / NT, k = NZ - 1;
        fw = fidx(d, 6, i, j, k); fa = fidx(d, 5, i, j, k); fb = fidx(d, 5, i, j, k - 1); fcc = fidx(d, 5, i, j, k - 2);
    }
    const float *W = v.wall;
    float S = W[(W_ZF + 0) * NB + b] * T[fa] - W[(W_ZF + 1) * NB + b] * T[fb] + W[(W_ZF + 2) * NB + b] * T[fcc];
    T[fw] = (S + rT[fw]) / W[(W_ZF + 3) * NB + b];
}

__global__ void __launch_bounds__(256) k_d_zero_flux(Dev d, View<float> v)
{
    if (d.ctl->done) return;
    d_zero_flux(d, v, blockIdx.x);
}

// outer update of the defect correction in FP64: x0 += dx, dx = 0, for the
// u, v, w, T cells and the u, v, w, p, T faces. With P_HALFRANGE_UPDATE it
// also reduces the largest p update and the largest negated one, each times
// inv_inner (per FP32 pass), into the slot the next cycle's check reads
__device__ __forceinline__ void mix_update(const Dev &d, const View<double> &x, const View<float> &dx, int lb,
                                           double inv_inner)
{
    int t = lb * blockDim.x + threadIdx.x;
    if (t < d.N)
        for (int q = 0; q < 4; q++) {
            size_t s = (size_t)q * d.N + t;
            x.sc[s] += (double)dx.sc[s];
            dx.sc[s] = 0.f;
        }
    double hi = -INFINITY, lo = -INFINITY;
    if (t < d.NF)
        for (int q = 0; q < 5; q++) {
            size_t s = (size_t)q * d.NF + t;
            const double v = (double)dx.fc[s];
            x.fc[s] += v;
            dx.fc[s] = 0.f;
            if (q == F_P) {
                hi = v * inv_inner;
                lo = -hi;
            }
        }
    if (d.ctl->p_measure != P_HALFRANGE_UPDATE) return;
    hi = wave_max(hi);
    lo = wave_max(lo);
    if ((threadIdx.x & 63) == 0 && hi > -INFINITY) {
        const int slot = d.ctl->iter & 1, cp = blockIdx.x & (ERR_COPIES - 1);
        atomicMax(d.errs + err_slot(slot, E_PHI, cp), order_key(hi));
        atomicMax(d.errs + err_slot(slot, E_PLO, cp), order_key(lo));
    }
}

__global__ void __launch_bounds__(256) k_mix_update(Dev d, View<double> x, View<float> dx, double inv_inner)
{
    if (d.ctl->done) return;
    mix_update(d, x, dx, blockIdx.x, inv_inner);
}

// FP32 copy of an FP64 block
__global__ void __launch_bounds__(256) k_cast(const double *in, float *out, long n)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        out[t] = (float)in[t];
}

// ---------------------------------------------------------------------
// iteration control
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_solve_reset(Dev d, double t0, double t1, double t2, double t3, double t4,
                                                    int max_iter, int min_iter, int p_measure)
{
    int l = threadIdx.x;
    for (int q = l; q < 2 * NERR * ERR_COPIES; q += blockDim.x) d.errs[(size_t)q * ERR_STRIDE] = 0ull;
    if (l == 0) {
        Ctl *c = d.ctl;
        c->done = 0;
        c->iter = 0;
        c->status = 0;
        c->max_iter = max_iter;
        c->min_iter = min_iter;
        c->p_measure = p_measure;
        c->tol[0] = t0; c->tol[1] = t1; c->tol[2] = t2; c->tol[3] = t3; c->tol[4] = t4;
        for (int q = 0; q < 5; q++) c->err[q] = 1.;
    }
}

// end of one outer iteration: reduce the increments, test convergence
// (the while condition of integrate_nse, nse.f90:129), clear the slot the
// next iteration uses
__device__ __forceinline__ void check(const Dev &d)
{
    Ctl *ctl = d.ctl;
    if (ctl->done) return;
    int l = threadIdx.x;  // one wave
    int s = ctl->iter & 1;
    double e[5];
    for (int q = 0; q < 5; q++) {
        e[q] = wave_max(__longlong_as_double((long long)d.errs[err_slot(s, q, l)]));
        d.errs[err_slot(s ^ 1, q, l)] = 0ull;
    }
    // largest pressure increment and largest negated one (keys: unsigned max)
    unsigned long long khi = d.errs[err_slot(s, E_PHI, l)], klo = d.errs[err_slot(s, E_PLO, l)];
    d.errs[err_slot(s ^ 1, E_PHI, l)] = 0ull;
    d.errs[err_slot(s ^ 1, E_PLO, l)] = 0ull;
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long a = __shfl_xor(khi, o), b = __shfl_xor(klo, o);
        khi = a > khi ? a : khi;
        klo = b > klo ? b : klo;
    }
    if (l == 0) {
        int it = ctl->iter + 1;
        ctl->iter = it;
        bool fin = true, conv = it >= ctl->min_iter;
        for (int q = 0; q < 5; q++) fin = fin && isfinite(e[q]);
        // half the range of the increments: blind to a change of p's
        // constant level (no pressure value this pass: the largest magnitude)
        if (ctl->p_measure != P_MAXABS && fin && khi && klo) e[3] = (order_value(khi) + order_value(klo)) / 2;
        for (int q = 0; q < 5; q++) {
            ctl->err[q] = e[q];
            conv = conv && !(e[q] > ctl->tol[q]);
        }
        if (!fin) { ctl->status = 4; ctl->done = 1; }
        else if (conv) { ctl->status = 0; ctl->done = 1; }
        else if (it >= ctl->max_iter) { ctl->status = 3; ctl->done = 1; }
    }
}

__global__ void __launch_bounds__(64) k_check(Dev d) { check(d); }

// ---------------------------------------------------------------------
// once-per-step and init kernels
// ---------------------------------------------------------------------
// calculate_pressure_nim_constants (nim.f90:556-743) at u = 0 (Taylor
// branch of B_nim..J_nim): C = -1/(3K), E = F = Det/2, G = I = 1/2, H = 1/K,
// J = -1/K, K = g^ii/rho
struct PCoef { double C[3], E[3], F[3], H[3], J[3], Det; };

__device__ PCoef pcoef(const Dev &d, int c)
{
    PCoef p;
    const int N = d.N;
    for (int q = 0; q < 3; q++) {
        double K = d.gc[(G_GD + q) * N + c] / d.rho;
        p.C[q] = -(1. / (3. * K));
        p.H[q] = 1. / K;
        p.J[q] = -(1. / K);
    }
    p.Det = 1. / (p.C[0] * p.C[1] + p.C[1] * p.C[2] + p.C[0] * p.C[2]);
    for (int q = 0; q < 3; q++) {
        p.E[q] = p.Det / 2.;
        p.F[q] = p.Det / 2.;
    }
    return p;
}

__global__ void __launch_bounds__(256) k_pressure_coeffs(Dev d)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.N) return;
    const int N = d.N;
    const double GG = 0.5, II = 0.5;
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    PCoef a = pcoef(d, c);
    for (int dir = 0; dir < 3; dir++) {
        int c2;
        if (dir == 0) { if (j == d.NR - 1) continue; c2 = c + d.NT; }
        else if (dir == 1) c2 = cidx(d, ip1(d, i), j, k);
        else { if (k == d.NZ - 1) continue; c2 = c + d.NT * d.NR; }
        PCoef b = pcoef(d, c2);
        int e1 = dir == 0 ? 1 : 0, e2 = dir == 2 ? 1 : 2;  // tangential directions (0-based)
        double gp = d.gc[(G_GF + 2 * dir + 1) * N + c], gm = d.gc[(G_GF + 2 * dir) * N + c2];
        double A[16];
        double Cea = a.C[e1] + a.C[e2], Ceb = b.C[e1] + b.C[e2];
        // normal pair 2dir+1, 2dir+2 (1-based slots), tangential pairs
        int n0 = 2 * dir + 1;
        A[n0] = -(GG + Cea * a.E[dir] * a.H[dir]) * gp;
        A[n0 + 1] = (GG - Cea * a.F[dir] * a.H[dir]) * gp + (II + Ceb * b.E[dir] * b.J[dir]) * gm;
        A[n0 + 2] = -(II - Ceb * b.F[dir] * b.J[dir]) * gm;
        // slot layout per direction (nim.f90:622-697): the 11 face slots are
        // ordered by face position 1..6 of cell and neighbour
        int te[2] = {e1, e2};
        double tv[2][4];
        for (int t = 0; t < 2; t++) {
            int oth = te[1 - t];
            tv[t][0] = a.C[oth] * a.E[te[t]] * a.H[dir] * gp;
            tv[t][1] = a.C[oth] * a.F[te[t]] * a.H[dir] * gp;
            tv[t][2] = -(b.C[oth] * b.E[te[t]] * b.J[dir] * gm);
            tv[t][3] = -(b.C[oth] * b.F[te[t]] * b.J[dir] * gm);
        }
        double *out = d.AA + (size_t)(dir * 15) * N;
        if (dir == 0) {
            // AA_px: 1 self-, 2 diag, 3 far, 4..7 theta pair (e1), 8..11 z pair (e2)
            out[0 * N + c] = A[1]; out[1 * N + c] = A[2]; out[2 * N + c] = A[3];
            out[3 * N + c] = tv[0][0]; out[4 * N + c] = tv[0][1]; out[5 * N + c] = tv[0][2]; out[6 * N + c] = tv[0][3];
            out[7 * N + c] = tv[1][0]; out[8 * N + c] = tv[1][1]; out[9 * N + c] = tv[1][2]; out[10 * N + c] = tv[1][3];
        } else if (dir == 1) {
            // AA_py: 1..4 r pair, 5 self-, 6 diag, 7 far, 8..11 z pair
            out[0 * N + c] = tv[0][0]; out[1 * N + c] = tv[0][1]; out[2 * N + c] = tv[0][2]; out[3 * N + c] = tv[0][3];
            out[4 * N + c] = A[3]; out[5 * N + c] = A[4]; out[6 * N + c] = A[5];
            out[7 * N + c] = tv[1][0]; out[8 * N + c] = tv[1][1]; out[9 * N + c] = tv[1][2]; out[10 * N + c] = tv[1][3];
        } else {
            // AA_pz: 1..4 r pair, 5..8 theta pair, 9 self-, 10 diag, 11 far
            out[0 * N + c] = tv[0][0]; out[1 * N + c] = tv[0][1]; out[2 * N + c] = tv[0][2]; out[3 * N + c] = tv[0][3];
            out[4 * N + c] = tv[1][0]; out[5 * N + c] = tv[1][1]; out[6 * N + c] = tv[1][2]; out[7 * N + c] = tv[1][3];
            out[8 * N + c] = A[5]; out[9 * N + c] = A[6]; out[10 * N + c] = A[7];
        }
        double s = a.C[e1] * a.C[e2] * a.Det * a.H[dir] * gp;
        double s2 = b.C[e1] * b.C[e2] * b.Det * b.J[dir] * gm;
        out[11 * N + c] = s;
        out[12 * N + c] = -s;
        out[13 * N + c] = -s2;
        out[14 * N + c] = s2;
        // the sweep's copy: tv[t][0] = tv[t][1] and tv[t][2] = tv[t][3]
        // (E = F), so one of each pair
        double *q = d.AQ + (size_t)(dir * NAQ) * N;
        q[Q_SELF * N + c] = A[n0];
        q[Q_DIAG * N + c] = A[n0 + 1];
        q[Q_FAR * N + c] = A[n0 + 2];
        q[Q_TA0 * N + c] = tv[0][0];
        q[Q_TB0 * N + c] = tv[0][2];
        q[Q_TA1 * N + c] = tv[1][0];
        q[Q_TB1 * N + c] = tv[1][2];
        q[Q_S1 * N + c] = -s;
        q[Q_S2 * N + c] = s2;
    }
}

// BCs_faces_ap (nim.f90:682-742) for every wall face
__global__ void __launch_bounds__(256) k_wall_pressure_coeffs(Dev d)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.NB) return;
    const int NT = d.NT, NR = d.NR, NZ = d.NZ, NB = d.NB, N = d.N;
    int i, j, k, pos;
    if (b < NT * NR) { pos = 5; i = b % NT; j = b / NT; k = 0; }
    else if (b < 2 * NT * NR) { pos = 6; i = (b - NT * NR) % NT; j = (b - NT * NR) / NT; k = NZ - 1; }
    else if (b < 2 * NT * NR + NT * NZ) { pos = 1; i = (b - 2 * NT * NR) % NT; k = (b - 2 * NT * NR) / NT; j = 0; }
    else { pos = 2; i = (b - 2 * NT * NR - NT * NZ) % NT; k = (b - 2 * NT * NR - NT * NZ) / NT; j = NR - 1; }
    int c = cidx(d, i, j, k);
    PCoef a = pcoef(d, c);
    int dir = (pos - 1) / 2;
    int e1 = dir == 0 ? 1 : 0, e2 = dir == 2 ? 1 : 2;
    bool plus = (pos % 2) == 0;
    double g = d.gc[(G_GF + pos - 1) * N + c];
    double L = plus ? a.H[dir] : a.J[dir];
    double Ce = a.C[e1] + a.C[e2];
    double ap[7];
    ap[2 * dir + 1] = -(0.5 + Ce * a.E[dir] * L) * g;
    ap[2 * dir + 2] = (0.5 - Ce * a.F[dir] * L) * g;
    ap[2 * e1 + 1] = a.C[e2] * a.E[e1] * L * g;
    ap[2 * e1 + 2] = a.C[e2] * a.F[e1] * L * g;
    ap[2 * e2 + 1] = a.C[e1] * a.E[e2] * L * g;
    ap[2 * e2 + 2] = a.C[e1] * a.F[e2] * L * g;
    double s7 = a.C[e1] * a.C[e2] * a.Det * L * g;
    for (int m = 1; m <= 6; m++) d.wall[(W_AP + m - 1) * NB + b] = ap[m];
    d.wall[(W_AP + 6) * NB + b] = s7;
    d.wall[(W_AP + 7) * NB + b] = -s7;
}

// update_time_step (nse.f90:1292-1313)
__device__ __forceinline__ void advance_cell(const Dev &d, int c)
{
    if (c >= d.N) return;
    const int N = d.N;
    double *s = d.sc;
    for (int q = 0; q < 4; q++) {
        double cur = s[(S_CUR + q) * N + c], old = s[(S_OLD + q) * N + c];
        s[(S_TC + q) * N + c] = (old + cur) / 2.;
        s[(S_OLD + q) * N + c] = cur;
    }
    int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
    const double *P = d.fc + (size_t)F_P * d.NF;
    int f[6];
    cell_faces(d, i, j, k, f);
    s[S_P * N + c] = (P[f[0]] + P[f[1]] + P[f[2]] + P[f[3]] + P[f[4]] + P[f[5]]) / 6.0;
    s[S_DILAO * N + c] = s[S_DILA * N + c];
}

__global__ void __launch_bounds__(256) k_advance(Dev d) { advance_cell(d, blockIdx.x * blockDim.x + threadIdx.x); }

// per-block partial results of check_steady_state (nse.f90:1316-1332) and
// evaluate_D (nse.f90:2504-2537): red[b*6 + {du, dv, dw, dT, sumD2}]
__device__ __forceinline__ void end_step_partial(const Dev &d, int lb)
{
    int c = lb * blockDim.x + threadIdx.x;
    double m[4] = {0., 0., 0., 0.}, s = 0.;
    if (c < d.N) {
        const int N = d.N;
        double den = 2 * d.DT;
        for (int q = 0; q < 4; q++) m[q] = fabs(d.sc[(S_CUR + q) * N + c] - d.sc[(S_OLD + q) * N + c]) / den;
        int i = iof(d, c % d.NT), j = (c / d.NT) % d.NR, k = c / (d.NT * d.NR);
        int f[6];
        cell_faces(d, i, j, k, f);
        double uf[6], vf[6], wf[6];
        for (int p = 0; p < 6; p++) {
            uf[p] = d.fc[f[p]];
            vf[p] = d.fc[d.NF + f[p]];
            wf[p] = d.fc[2 * (size_t)d.NF + f[p]];
        }
        double invV = d.gc[G_INVV * N + c];
        double D = dphi(d, d.An, uf, f, 0, invV) + dphi(d, d.An, vf, f, 1, invV) + dphi(d, d.An, wf, f, 2, invV);
        d.sc[S_DILA * N + c] = D;
        s = D * D;
    }
    __shared__ double sh[4][5];
    int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < 4; q++) m[q] = wave_max(m[q]);
    s = wave_sum(s);
    if (l == 0) {
        for (int q = 0; q < 4; q++) sh[wv][q] = m[q];
        sh[wv][4] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 5; q++) {
            double v = sh[0][q];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = q < 4 ? fmax(v, sh[w][q]) : v + sh[w][q];
            d.red[(size_t)lb * 6 + q] = v;
        }
    }
}

__global__ void __launch_bounds__(256) k_end_step_partial(Dev d) { end_step_partial(d, blockIdx.x); }

// combine the partials in a fixed order; out = {dudt_max, dTdt_max, Dilatation}
__device__ __forceinline__ void end_step_final(const Dev &d, int nblocks, double *out)
{
    int l = threadIdx.x;  // 256 threads
    double m[4] = {0., 0., 0., 0.}, s = 0.;
    for (int b = l; b < nblocks; b += blockDim.x) {
        for (int q = 0; q < 4; q++) m[q] = fmax(m[q], d.red[(size_t)b * 6 + q]);
        s += d.red[(size_t)b * 6 + 4];
    }
    __shared__ double sh[4][5];
    int wv = l >> 6;
    for (int q = 0; q < 4; q++) m[q] = wave_max(m[q]);
    s = wave_sum(s);
    if ((l & 63) == 0) {
        for (int q = 0; q < 4; q++) sh[wv][q] = m[q];
        sh[wv][4] = s;
    }
    __syncthreads();
    if (l == 0) {
        double r[5];
        for (int q = 0; q < 5; q++) {
            r[q] = sh[0][q];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) r[q] = q < 4 ? fmax(r[q], sh[w][q]) : r[q] + sh[w][q];
        }
        out[0] = fmax(r[2], fmax(r[0], r[1]));
        out[1] = r[3];
        out[2] = sqrt(r[4] / d.N);
    }
}

__global__ void __launch_bounds__(256) k_end_step_final(Dev d, int nblocks, double *out)
{
    end_step_final(d, nblocks, out);
}

// find_Nu2 (nse.f90:2699-2798): one block, fixed summation order.
// out = {Nu_H, Nu_C, Average, Area1, Area2}
__device__ __forceinline__ void nusselt(const Dev &d, double R_i, double R_o, double *out)
{
    const int NT = d.NT, NR = d.NR, NZ = d.NZ, NB = d.NB;
    const double *T = d.fc + (size_t)F_T * d.NF, *W = d.wall;
    int n = NT * NZ;
    double s1 = 0., a1 = 0., s2 = 0., a2 = 0.;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        int i = q % NT, k = q / NT;
        // hexagon wall, position 1
        int b = 2 * NT * NR + q;
        double dd = -W[(W_NU + 0) * NB + b] * T[fidx(d, 1, i, 0, k)] + W[(W_NU + 1) * NB + b] * T[fidx(d, 2, i, 0, k)]
                    - W[(W_NU + 2) * NB + b] * T[fidx(d, 2, i, 1, k)] + W[(W_NU + 3) * NB + b] * T[fidx(d, 2, i, 2, k)];
        s2 += W[(W_NU + 4) * NB + b] * (-dd);
        a2 += W[(W_NU + 4) * NB + b];
        // cylinder wall, position 2
        b = 2 * NT * NR + NT * NZ + q;
        dd = -W[(W_NU + 0) * NB + b] * T[fidx(d, 2, i, NR - 1, k)] + W[(W_NU + 1) * NB + b] * T[fidx(d, 1, i, NR - 1, k)]
             - W[(W_NU + 2) * NB + b] * T[fidx(d, 1, i, NR - 2, k)] + W[(W_NU + 3) * NB + b] * T[fidx(d, 1, i, NR - 3, k)];
        s1 += W[(W_NU + 4) * NB + b] * dd;
        a1 += W[(W_NU + 4) * NB + b];
    }
    __shared__ double sh[4][4];
    int wv = threadIdx.x >> 6;
    s1 = wave_sum(s1); a1 = wave_sum(a1); s2 = wave_sum(s2); a2 = wave_sum(a2);
    if ((threadIdx.x & 63) == 0) { sh[wv][0] = s1; sh[wv][1] = a1; sh[wv][2] = s2; sh[wv][3] = a2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        s1 = sh[0][0]; a1 = sh[0][1]; s2 = sh[0][2]; a2 = sh[0][3];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            s1 += sh[w][0]; a1 += sh[w][1]; s2 += sh[w][2]; a2 += sh[w][3];
        }
        out[0] = -(s1 / a1 * R_o);
        out[1] = -(s2 / a2 * R_i);
        out[2] = -((s1 / a1 * R_o) + (s2 / a2 * R_i)) / 2.;
        out[3] = a1;
        out[4] = a2;
    }
}

__global__ void __launch_bounds__(256) k_nusselt(Dev d, double R_i, double R_o, double *out)
{
    nusselt(d, R_i, R_o, out);
}


// ---------------------------------------------------------------------
// ensembles: each kernel of an outer pass launched once over all cases
// ---------------------------------------------------------------------
// Block b (after the XCD renumbering) belongs to the case q with
// bstart[q] <= b < bstart[q + 1] and is block b - bstart[q] of that case's
// own launch, so every case runs exactly the blocks it runs alone.
__device__ __forceinline__ int ens_case(const Ens &e, int &lb)
{
    int b = xcd_block();
    int lo = 0, hi = e.ncase - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (e.bstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    lb = b - e.bstart[lo];
    return lo;
}

// a full-domain colour phase as a one-item schedule
__device__ __forceinline__ Sched full_phase(const Dev &d, int color)
{
    Sched sc{};
    sc.n = 1;
    sc.color[0] = color;
    sc.k0[0] = 0;
    sc.k1[0] = d.NZ;
    return sc;
}

#define ENS_CASE                                                                                        \
    int lb;                                                                                             \
    const int q = ens_case(e, lb);                                                                      \
    if (lb >= e.bstart[q + 1] - e.bstart[q]) return;                                                    \
    const Dev &d = e.dev[q];                                                                            \
    if (d.ctl->done) return;

__global__ void __launch_bounds__(256) k_e_mom_src(Ens e) { ENS_CASE mom_src_cell(d, view64(d), lb); }
__global__ void __launch_bounds__(256) k_e_mom_sweep(Ens e, int color)
{
    ENS_CASE mom_sweep_cell<MODE_GS>(d, view64(d), full_phase(d, color), lb);
}
__global__ void __launch_bounds__(256) k_e_ppe_src(Ens e) { ENS_CASE ppe_src<MODE_GS>(d, view64(d), lb); }
__global__ void __launch_bounds__(256) k_e_ppe_sweep(Ens e, int color)
{
    ENS_CASE ppe_sweep<MODE_GS, double, true>(d, view64(d), full_phase(d, color), lb);
}
__global__ void __launch_bounds__(256) k_e_T_coef(Ens e) { ENS_CASE T_coef<MODE_GS>(d, view64(d), lb); }
__global__ void __launch_bounds__(256) k_e_T_src(Ens e) { ENS_CASE T_src(d, view64(d), lb); }
__global__ void __launch_bounds__(256) k_e_T_sweep(Ens e, int color)
{
    ENS_CASE View<double> v = view64(d);
    T_sweep<MODE_GS, double, true>(d, v, v.TF, full_phase(d, color), lb);
}
// one block per case
__global__ void __launch_bounds__(64) k_e_check(Ens e) { check(e.dev[blockIdx.x]); }

__global__ void __launch_bounds__(192) k_e_res_mom(Ens e) { ENS_CASE mom_sweep<MODE_RES>(d, e.v64[q], Sched{}, lb); }
__global__ void __launch_bounds__(256) k_e_res_ppe_src(Ens e) { ENS_CASE ppe_src<MODE_RES>(d, e.v64[q], lb); }
__global__ void __launch_bounds__(256) k_e_res_ppe(Ens e) { ENS_CASE ppe_sweep<MODE_RES>(d, e.v64[q], Sched{}, lb); }
__global__ void __launch_bounds__(256) k_e_res_T_coef(Ens e) { ENS_CASE T_coef<MODE_RES>(d, e.v64[q], lb); }
__global__ void __launch_bounds__(256) k_e_res_T(Ens e)
{
    ENS_CASE T_sweep<MODE_RES>(d, e.v64[q], e.v64[q].TF, Sched{}, lb);
}
__global__ void __launch_bounds__(256) k_e_d_mom_src(Ens e) { ENS_CASE mom_src_cell(d, e.v32[q], lb); }
__global__ void __launch_bounds__(192) k_e_d_mom_sweep(Ens e, int color)
{
    ENS_CASE mom_sweep<MODE_DELTA>(d, e.v32[q], full_phase(d, color), lb);
}
__global__ void __launch_bounds__(256) k_e_d_ppe_src(Ens e) { ENS_CASE ppe_src<MODE_DELTA>(d, e.v32[q], lb); }
__global__ void __launch_bounds__(256) k_e_d_ppe_sweep(Ens e, int color)
{
    ENS_CASE ppe_sweep<MODE_DELTA>(d, e.v32[q], full_phase(d, color), lb);
}
__global__ void __launch_bounds__(256) k_e_d_zero_flux(Ens e) { ENS_CASE d_zero_flux(d, e.v32[q], lb); }
__global__ void __launch_bounds__(256) k_e_d_T_src(Ens e) { ENS_CASE T_src(d, e.v32[q], lb); }
__global__ void __launch_bounds__(256) k_e_d_T_sweep(Ens e, int color)
{
    ENS_CASE T_sweep<MODE_DELTA>(d, e.v32[q], e.v32[q].TF, full_phase(d, color), lb);
}
__global__ void __launch_bounds__(256) k_e_mix_update(Ens e, double inv_inner)
{
    ENS_CASE mix_update(d, e.v64[q], e.v32[q], lb, inv_inner);
}

// per-step work of every case (no early exit on convergence: it runs
// around the solve)
#define ENS_STEP_CASE                                                                                   \
    int lb;                                                                                             \
    const int q = ens_case(e, lb);                                                                      \
    if (lb >= e.bstart[q + 1] - e.bstart[q]) return;                                                    \
    const Dev &d = e.dev[q];

// momentum coefficients (calculate_velocity_nim_constants_formulation2),
// with the FP32 copy of the mixed loop when cast is set
__global__ void __launch_bounds__(256) k_e_coef_A(Ens e)
{
    ENS_STEP_CASE
    const int c = lb * blockDim.x + threadIdx.x;
    if (c < d.N) coef_A_cell(d, 0, c);
}
__global__ void __launch_bounds__(256) k_e_coef_F(Ens e, int cast)
{
    ENS_STEP_CASE
    const int c = lb * blockDim.x + threadIdx.x;
    double eT = 0.;
    coef_F_cell(d, 0, c, eT);
    if (cast && c < d.N)
        for (int m = 0; m < NCF; m++)
            const_cast<float *>(e.v32[q].MF)[(size_t)m * d.N + c] = (float)d.MF[(size_t)m * d.N + c];
}
// check_steady_state + evaluate_D partials (per-cell launch shape); then one
// block per case reduces them into out[8 q + 0..2]
__global__ void __launch_bounds__(256) k_e_end_step_partial(Ens e) { ENS_STEP_CASE end_step_partial(d, lb); }
__global__ void __launch_bounds__(256) k_e_end_step_final(Ens e, const int *nred, double *out)
{
    end_step_final(e.dev[blockIdx.x], nred[blockIdx.x], out + 8 * (size_t)blockIdx.x);
}
__global__ void __launch_bounds__(256) k_e_advance(Ens e)
{
    ENS_STEP_CASE
    advance_cell(d, lb * blockDim.x + threadIdx.x);
}
// find_Nu2 of every case, one block per case, into out[8 q + 3..7]
__global__ void __launch_bounds__(256) k_e_nusselt(Ens e, const double *R, double *out)
{
    nusselt(e.dev[blockIdx.x], R[2 * blockIdx.x], R[2 * blockIdx.x + 1], out + 8 * (size_t)blockIdx.x + 3);
}

// the iteration control of every case into one array (one host read)
__global__ void __launch_bounds__(64) k_e_gather(Ens e, Ctl *out)
{
    for (int q = threadIdx.x; q < e.ncase; q += blockDim.x) out[q] = *e.dev[q].ctl;
}

}  // namespace nim
