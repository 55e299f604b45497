/* CPU oracle for the NIM time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is a plain-C restatement of the
 * reference's per-time-step algorithm (ijarrah2/NIM-CONV22-SHARED,
 * src/nim.f90, src/nse.f90, src/bc.f90). Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg may load it, and only as the checker; the
 * product path never calls it.
 *
 * Data layout: the reference's own global arrays (parameters.f90:222-292)
 * with 0-based indices, Fortran column-major inner dimensions kept, i.e.
 * xi_x(a,b,c) -> xi_x[c*9 + (a-1) + 3*(b-1)], gij_faces(a,b,f,c) ->
 * gij_faces[c*54 + (a-1) + 3*(b-1) + 9*(f-1)], M_F4(m,c) -> M_F4[c*8+m-1].
 * Face and cell numbers follow the reference numbering (create_geometry.f90
 * :321-664), so arrays dumped from the compiled reference load unchanged.
 * Wall neighbours are -1 (the reference uses 0 in 1-based numbering).
 *
 * Two sweep orders are provided:
 *  - ORA_NATURAL: in-place Gauss-Seidel in natural cell order with the
 *    boundary updates as separate passes, exactly as the reference runs with
 *    OMP_NUM_THREADS=1 (run.sh:2). Expression order follows the Fortran
 *    source so results agree with the compiled reference to rounding.
 *  - ORA_COLORED: the race-free order of the GPU path. Cells get colour
 *    (2*i_theta + j_r + 3*k_z) mod 4; the four colours are swept in turn and
 *    each cell does the same per-cell work as the reference (cell value, then
 *    owned faces 2, 4, 6). In the pressure stage each cell first updates its
 *    own wall faces (bottom, top, hexagon, cylinder, the reference's boundary
 *    group order) and then its interior faces. Convergence is measured as the
 *    maximum change since the start of the solve call.
 *
 * Parity status of each function is pinned by tests/test_oracle_reference.py
 * (against the compiled reference's dumps) and tests/test_oracle_math.py
 * (closed forms at high precision, limits, both sweep orders).
 *
 * Loops whose iterations write disjoint values and read none of them (the
 * sources, the constants, each colour phase of the colored order, the
 * error maxima) run under OpenMP when built with it; the natural-order
 * sweeps and boundary passes stay serial. Results therefore do not depend
 * on the thread count.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* meshes smaller than this run the loops on one thread: on a few thousand
 * cells a parallel region costs more than its loop, and several test
 * processes sharing the host's cores make spinning teams far slower still */
#define ORA_PAR_MIN 16384
#endif

enum { ORA_NATURAL = 0, ORA_COLORED = 1 };

typedef struct {
    double nu, Re, alpha, rho, g_beta, DT;
    double w_u, w_v, w_w, w_p, w_p2, w_T;
    double v_limit, R_i, R_o;
} ora_params;

typedef struct {
    int nt, nr, nz; /* O-grid cells in theta, r, z */
    int nmesh, nface, nbc;
    ora_params prm;
    double TI;
    int STEP;
    /* topology (0-based) */
    int32_t *mfaces, *nbr, *faces_bc, *bc_faces, *bc_pos, *bc_mesh, *bc_points;
    /* geometry */
    double *volume, *xi_x, *gij, *gij_faces, *G1, *G2, *G3;
    double *faces_area, *faces_normal, *faces_mid;
    double *bc_a, *bc_ap, *AA_px, *AA_py, *AA_pz;
    /* state */
    double *faces_u, *faces_v, *faces_w, *faces_p, *faces_T;
    double *faces_u_old, *faces_v_old, *faces_w_old, *faces_p_old, *faces_T_old;
    double *uxyz, *vxyz, *wxyz, *Txyz;
    double *uxyz_old, *vxyz_old, *wxyz_old, *Txyz_old;
    double *uxyzt, *vxyzt, *wxyzt, *Txyzt;
    double *uxyz_old_con, *vxyz_old_con, *wxyz_old_con, *Txyz_old_con;
    double *meshes_p, *dila, *dila_old;
    /* per-step and per-iteration work arrays */
    double *meshes_u, *meshes_v, *meshes_w;
    double *M_A, *M_F1, *M_F2, *M_F3, *M_F4;
    double *MT_A, *MT_F1, *MT_F2, *MT_F3, *MT_F4;
    double *meshes_bz, *faces_bz;
    double *R_B_u, *R_B_v, *R_B_w, *R_C_u, *R_C_v, *R_C_w, *R_D_u, *R_D_v, *R_D_w;
    double *R_B_p, *R_D_p, *R_D_T;
    double Dilatation, dudt_max, dTdt_max;
    /* p's convergence measure: 0 max |dp| (find_error, nse.f90:260-325);
     * 1 half the range of dp, (max dp - min dp) / 2, which does not see a
     * change of p's constant level (p is fixed only up to a constant) */
    int p_measure;
} ora_ctx;

/* ------------------------------------------------------------------ */
/* named array table for the Python wrapper                            */
/* ------------------------------------------------------------------ */
enum { K_MESH, K_FACE, K_BC };
typedef struct { const char *name; size_t off; int kind; int mult; int is_int; } ora_field;
#define F(nm, k, m) { #nm, offsetof(ora_ctx, nm), k, m, 0 }
#define I(nm, k, m) { #nm, offsetof(ora_ctx, nm), k, m, 1 }
static const ora_field FIELDS[] = {
    I(mfaces, K_MESH, 6), I(nbr, K_MESH, 6), I(faces_bc, K_FACE, 1), I(bc_faces, K_BC, 1),
    I(bc_pos, K_BC, 1), I(bc_mesh, K_BC, 1), I(bc_points, K_BC, 5),
    F(volume, K_MESH, 1), F(xi_x, K_MESH, 9), F(gij, K_MESH, 9), F(gij_faces, K_MESH, 54),
    F(G1, K_MESH, 1), F(G2, K_MESH, 1), F(G3, K_MESH, 1),
    F(faces_area, K_FACE, 1), F(faces_normal, K_FACE, 3), F(faces_mid, K_FACE, 3),
    F(bc_a, K_BC, 20), F(bc_ap, K_BC, 8), F(AA_px, K_MESH, 15), F(AA_py, K_MESH, 15), F(AA_pz, K_MESH, 15),
    F(faces_u, K_FACE, 1), F(faces_v, K_FACE, 1), F(faces_w, K_FACE, 1), F(faces_p, K_FACE, 1), F(faces_T, K_FACE, 1),
    F(faces_u_old, K_FACE, 1), F(faces_v_old, K_FACE, 1), F(faces_w_old, K_FACE, 1), F(faces_p_old, K_FACE, 1),
    F(faces_T_old, K_FACE, 1),
    F(uxyz, K_MESH, 1), F(vxyz, K_MESH, 1), F(wxyz, K_MESH, 1), F(Txyz, K_MESH, 1),
    F(uxyz_old, K_MESH, 1), F(vxyz_old, K_MESH, 1), F(wxyz_old, K_MESH, 1), F(Txyz_old, K_MESH, 1),
    F(uxyzt, K_MESH, 1), F(vxyzt, K_MESH, 1), F(wxyzt, K_MESH, 1), F(Txyzt, K_MESH, 1),
    F(uxyz_old_con, K_MESH, 1), F(vxyz_old_con, K_MESH, 1), F(wxyz_old_con, K_MESH, 1), F(Txyz_old_con, K_MESH, 1),
    F(meshes_p, K_MESH, 1), F(dila, K_MESH, 1), F(dila_old, K_MESH, 1),
    F(meshes_u, K_MESH, 1), F(meshes_v, K_MESH, 1), F(meshes_w, K_MESH, 1),
    F(M_A, K_MESH, 18), F(M_F1, K_MESH, 5), F(M_F2, K_MESH, 5), F(M_F3, K_MESH, 5), F(M_F4, K_MESH, 8),
    F(MT_A, K_MESH, 18), F(MT_F1, K_MESH, 5), F(MT_F2, K_MESH, 5), F(MT_F3, K_MESH, 5), F(MT_F4, K_MESH, 8),
    F(meshes_bz, K_MESH, 1), F(faces_bz, K_FACE, 1),
    F(R_B_u, K_MESH, 1), F(R_B_v, K_MESH, 1), F(R_B_w, K_MESH, 1), F(R_C_u, K_MESH, 1), F(R_C_v, K_MESH, 1),
    F(R_C_w, K_MESH, 1), F(R_D_u, K_MESH, 1), F(R_D_v, K_MESH, 1), F(R_D_w, K_MESH, 1),
    F(R_B_p, K_MESH, 1), F(R_D_p, K_MESH, 1), F(R_D_T, K_MESH, 1),
};
#undef F
#undef I
#define NFIELDS ((int)(sizeof(FIELDS) / sizeof(FIELDS[0])))

static size_t field_len(const ora_ctx *c, const ora_field *f)
{
    size_t n = f->kind == K_MESH ? (size_t)c->nmesh : f->kind == K_FACE ? (size_t)c->nface : (size_t)c->nbc;
    return n * (size_t)f->mult;
}

ora_ctx *ora_create(int nt, int nr, int nz, const ora_params *prm)
{
    ora_ctx *c = (ora_ctx *)calloc(1, sizeof(ora_ctx));
    if (!c) return NULL;
    c->nt = nt; c->nr = nr; c->nz = nz;
    c->nmesh = nt * nr * nz;
    c->nface = 3 * nt * nr * nz + nt * nz + nt * nr; /* parameters.f90:117 */
    c->nbc = 2 * nt * nr + 2 * nt * nz;
    c->prm = *prm;
    c->TI = prm->DT; /* nse.f90:31-32 */
    c->STEP = 1;
    for (int i = 0; i < NFIELDS; i++) {
        size_t n = field_len(c, &FIELDS[i]);
        void *p = calloc(n ? n : 1, FIELDS[i].is_int ? sizeof(int32_t) : sizeof(double));
        if (!p) return NULL;
        *(void **)((char *)c + FIELDS[i].off) = p;
    }
    return c;
}

void ora_destroy(ora_ctx *c)
{
    if (!c) return;
    for (int i = 0; i < NFIELDS; i++) free(*(void **)((char *)c + FIELDS[i].off));
    free(c);
}

/* returns element count and type (1 int32, 0 float64), or -1 if unknown */
long ora_field_info(ora_ctx *c, const char *name, int *is_int)
{
    for (int i = 0; i < NFIELDS; i++)
        if (!strcmp(FIELDS[i].name, name)) {
            *is_int = FIELDS[i].is_int;
            return (long)field_len(c, &FIELDS[i]);
        }
    return -1;
}

void *ora_field_ptr(ora_ctx *c, const char *name)
{
    for (int i = 0; i < NFIELDS; i++)
        if (!strcmp(FIELDS[i].name, name)) return *(void **)((char *)c + FIELDS[i].off);
    return NULL;
}

void ora_get_scalars(ora_ctx *c, double out[6])
{
    out[0] = c->Dilatation; out[1] = c->dudt_max; out[2] = c->dTdt_max;
    out[3] = c->TI; out[4] = c->STEP; out[5] = 0;
}

void ora_set_time(ora_ctx *c, double TI, int STEP) { c->TI = TI; c->STEP = STEP; }

/* threads the OpenMP loops use (1 without OpenMP) */
int ora_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ora_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* index helpers (1-based position/component arguments as in Fortran)  */
/* ------------------------------------------------------------------ */
#define MF(c_, n, f) ((c_)->mfaces[(size_t)(n) * 6 + (f) - 1])
#define NB(c_, n, f) ((c_)->nbr[(size_t)(n) * 6 + (f) - 1])
#define XI(c_, n, a, b) ((c_)->xi_x[(size_t)(n) * 9 + (a) - 1 + 3 * ((b) - 1)])
#define GIJ(c_, n, a, b) ((c_)->gij[(size_t)(n) * 9 + (a) - 1 + 3 * ((b) - 1)])
#define GF(c_, n, a, b, f) ((c_)->gij_faces[(size_t)(n) * 54 + (a) - 1 + 3 * ((b) - 1) + 9 * ((f) - 1)])
#define NRM(c_, f, d) ((c_)->faces_normal[(size_t)(f) * 3 + (d) - 1])
#define MID(c_, f, d) ((c_)->faces_mid[(size_t)(f) * 3 + (d) - 1])
#define AREA(c_, f) ((c_)->faces_area[f])
#define A3(arr, n, d, m) ((arr)[(size_t)(n) * 18 + (d) - 1 + 3 * ((m) - 1)])
#define F5(arr, n, m) ((arr)[(size_t)(n) * 5 + (m) - 1])
#define F8(arr, n, m) ((arr)[(size_t)(n) * 8 + (m) - 1])
#define A15(arr, n, m) ((arr)[(size_t)(n) * 15 + (m) - 1])
#define BCA(c_, i, m) ((c_)->bc_a[(size_t)(i) * 20 + (m) - 1])
#define BCAP(c_, i, m) ((c_)->bc_ap[(size_t)(i) * 8 + (m) - 1])
#define BCPT(c_, i, m) ((c_)->bc_points[(size_t)(i) * 5 + (m) - 1])

/* x**n for a constant integer n, by the binary method that LLVM uses to
 * expand powi (the reference's integer powers, e.g. k11**10, nim.f90:197) */
static double ipow(double x, int n)
{
    double r = 1.0, cur = x;
    int first = 1;
    while (n) {
        if (n & 1) { r = first ? cur : r * cur; first = 0; }
        n >>= 1;
        if (n) cur = cur * cur;
    }
    return r;
}

static int cell_color(const ora_ctx *c, int n)
{
    int i = n % c->nt, j = (n / c->nt) % c->nr, k = n / (c->nt * c->nr);
    return (2 * i + j + 3 * k) & 3;
}

/* ------------------------------------------------------------------ */
/* NIM coefficients, one direction (nim.f90:184-287; a = 1)            */
/* ------------------------------------------------------------------ */
static void nim_direction(double u, double k, double gp, double gm, double vlim, double A[6])
{
    const double a = 1.0;
    double Re = 2 * a * u / k;
    if (fabs(Re) > vlim) {
        A[0] = (2 * a / (k * (1. / exp(Re) - 1.)) + 1. / u) * gp;
        A[1] = (-(2 * a / (k * (1. - exp(Re)))) - 1. / u) * gm;
        A[2] = (u / (k * (1. / exp(Re) - 1.))) * gp;
        A[3] = (u / (k * (1. - exp(Re)))) * gm;
        if (Re > 0) A[4] = (-k + (a * u * (1. / exp(Re) + 1.) / (-1. / exp(Re) + 1.))) / (u * u);
        if (Re < 0) A[4] = (-k + (a * u * (1. + exp(Re)) / (-1. + exp(Re)))) / (u * u);
        if (Re > 0) A[5] = 1. / (-1. / exp(Re) + 1.) - 1. / Re;
        if (Re < 0) A[5] = exp(Re) / (-1. + exp(Re)) - 1. / Re;
    } else {
        double u2 = ipow(u, 2), u3 = ipow(u, 3), u4 = ipow(u, 4), u5 = ipow(u, 5), u6 = ipow(u, 6);
        double u7 = ipow(u, 7), u8 = ipow(u, 8), u9 = ipow(u, 9), u10 = ipow(u, 10);
        double k2 = ipow(k, 2), k3 = ipow(k, 3), k4 = ipow(k, 4), k5 = ipow(k, 5), k6 = ipow(k, 6);
        double k7 = ipow(k, 7), k8 = ipow(k, 8), k9 = ipow(k, 9), k10 = ipow(k, 10), k11 = ipow(k, 11);
        A[0] = (-(a / k) - u / (3. * k2) + u3 / (45. * k4) - 2 * u5 / (945. * k6) + -(u7 / (4725. * k8))
                - 2 * u9 / (93555. * k10)) * gp;
        A[1] = (-(a / k) + u / (3. * k2) - u3 / (45. * k4) + 2 * u5 / (945. * k6) - -(u7 / (4725. * k8))
                + 2 * u9 / (93555. * k10)) * gm;
        A[2] = (-(1 / (2. * a)) - u / (2. * k) - u2 / (6. * k2) + u4 / (90. * k4) - -(u6 / (945. * k6))
                + u8 / (9450. * k8) - u10 / (93555. * k10)) * gp;
        A[3] = (-(1 / (2. * a)) + u / (2. * k) - u2 / (6. * k2) + u4 / (90. * k4) - -(u6 / (945. * k6))
                + u8 / (9450. * k8) - u10 / (93555. * k10)) * gm;
        A[4] = a / (3. * k) - u2 / (45. * k3) + 2 * u4 / (945. * k5) - -(u6 / (4725. * k7))
               + 2 * u8 / (93555. * k9) - -(1382 * u10 / (6.38512875e8 * k11));
        A[5] = 0.5 + u / (6. * k) - u3 / (90. * k3) + u5 / (945. * k5) - -(u7 / (9450. * k7)) + u9 / (93555. * k9);
    }
}

/* the six constants of one direction, for tests of the closed forms */
void ora_nim_direction(double u, double k, double gp, double gm, double vlim, double A[6])
{
    nim_direction(u, k, gp, gm, vlim, A);
}

/* F4, F1..F3 from the exponential-fit constants (nim.f90:292-329) */
static void nim_face_coeffs(ora_ctx *c, const double *A, double *F1, double *F2, double *F3, double *F4)
{
    const double DT = c->prm.DT;
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int i = 0; i < c->nmesh; i++) {
        F8(F4, i, 2) = 0.5 * (1. / A3(A, i, 1, 5) + 1. / A3(A, i, 2, 5) + 1. / A3(A, i, 3, 5) - 1. / DT);
        F8(F4, i, 3) = -(A3(A, i, 1, 6) / A3(A, i, 1, 5));
        F8(F4, i, 4) = (A3(A, i, 1, 6) - 1.) / A3(A, i, 1, 5);
        F8(F4, i, 5) = -(A3(A, i, 2, 6) / A3(A, i, 2, 5));
        F8(F4, i, 6) = (A3(A, i, 2, 6) - 1.) / A3(A, i, 2, 5);
        F8(F4, i, 7) = -(A3(A, i, 3, 6) / A3(A, i, 3, 5));
        F8(F4, i, 8) = (A3(A, i, 3, 6) - 1.) / A3(A, i, 3, 5);
        F8(F4, i, 1) = F8(F4, i, 2) + F8(F4, i, 3) + F8(F4, i, 4) + F8(F4, i, 5) + F8(F4, i, 6) + F8(F4, i, 7)
                       + F8(F4, i, 8);
        double *FF[3] = {F1, F2, F3};
        const int pos[3] = {2, 4, 6};
        for (int d = 1; d <= 3; d++) {
            int i2 = NB(c, i, pos[d - 1]);
            if (i2 < 0) continue;
            double *Fd = FF[d - 1];
            F5(Fd, i, 2) = A3(A, i, d, 3) - (A3(A, i, d, 1) * A3(A, i, d, 6) / A3(A, i, d, 5));
            F5(Fd, i, 3) = A3(A, i2, d, 4) - A3(A, i2, d, 2) / A3(A, i2, d, 5)
                           + A3(A, i2, d, 2) * A3(A, i2, d, 6) / A3(A, i2, d, 5);
            F5(Fd, i, 4) = A3(A, i, d, 1) / (2. * A3(A, i, d, 5));
            F5(Fd, i, 5) = A3(A, i2, d, 2) / (2. * A3(A, i2, d, 5));
            F5(Fd, i, 1) = F5(Fd, i, 2) + F5(Fd, i, 3) + 2. * F5(Fd, i, 4) + 2. * F5(Fd, i, 5);
        }
    }
}

/* calculate_velocity_nim_constants_formulation2 (nim.f90:160-330) */
void ora_velocity_constants(ora_ctx *c)
{
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int i = 0; i < c->nmesh; i++) {
        c->meshes_u[i] = c->uxyzt[i] * XI(c, i, 1, 1) + c->vxyzt[i] * XI(c, i, 1, 2) + c->wxyzt[i] * XI(c, i, 1, 3);
        c->meshes_v[i] = c->uxyzt[i] * XI(c, i, 2, 1) + c->vxyzt[i] * XI(c, i, 2, 2) + c->wxyzt[i] * XI(c, i, 2, 3);
        c->meshes_w[i] = c->uxyzt[i] * XI(c, i, 3, 1) + c->vxyzt[i] * XI(c, i, 3, 2) + c->wxyzt[i] * XI(c, i, 3, 3);
        const double vel[3] = {c->meshes_u[i], c->meshes_v[i], c->meshes_w[i]};
        for (int d = 1; d <= 3; d++) {
            double A[6];
            nim_direction(vel[d - 1], GIJ(c, i, d, d) * c->prm.nu, GF(c, i, d, d, 2 * d), GF(c, i, d, d, 2 * d - 1),
                          c->prm.v_limit, A);
            for (int m = 1; m <= 6; m++) A3(c->M_A, i, d, m) = A[m - 1];
        }
    }
    nim_face_coeffs(c, c->M_A, c->M_F1, c->M_F2, c->M_F3, c->M_F4);
}

/* calculate_temperature_nim_constants_formulation2 (nim.f90:334-503) */
void ora_temperature_constants(ora_ctx *c)
{
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int i = 0; i < c->nmesh; i++) {
        double ut = 0.5 * (c->uxyz[i] + c->uxyz_old[i]);
        double vt = 0.5 * (c->vxyz[i] + c->vxyz_old[i]);
        double wt = 0.5 * (c->wxyz[i] + c->wxyz_old[i]);
        const double vel[3] = {ut * XI(c, i, 1, 1) + vt * XI(c, i, 1, 2) + wt * XI(c, i, 1, 3),
                               ut * XI(c, i, 2, 1) + vt * XI(c, i, 2, 2) + wt * XI(c, i, 2, 3),
                               ut * XI(c, i, 3, 1) + vt * XI(c, i, 3, 2) + wt * XI(c, i, 3, 3)};
        for (int d = 1; d <= 3; d++) {
            double A[6];
            nim_direction(vel[d - 1], GIJ(c, i, d, d) * c->prm.alpha, GF(c, i, d, d, 2 * d),
                          GF(c, i, d, d, 2 * d - 1), c->prm.v_limit, A);
            for (int m = 1; m <= 6; m++) A3(c->MT_A, i, d, m) = A[m - 1];
        }
    }
    nim_face_coeffs(c, c->MT_A, c->MT_F1, c->MT_F2, c->MT_F3, c->MT_F4);
}

/* calculate_pressure_nim_constants (nim.f90:556-743). The pressure equation
 * is pure diffusion (u = 0), so B_nim..J_nim (nim.f90:757-875) take their
 * Taylor branch at u = 0 and reduce to C = -1/(3K), E = F = Det/2, G = I = 1/2,
 * H = 1/K, J = -1/K. */
void ora_pressure_constants(ora_ctx *c)
{
    const int N = c->nmesh;
    double *CC = (double *)malloc(sizeof(double) * 3 * N), *EE = (double *)malloc(sizeof(double) * 3 * N);
    double *FF = (double *)malloc(sizeof(double) * 3 * N), *HH = (double *)malloc(sizeof(double) * 3 * N);
    double *JJ = (double *)malloc(sizeof(double) * 3 * N), *Det = (double *)malloc(sizeof(double) * N);
    const double GG = 0.5, II = 0.5;
#define C3(arr, d, n) ((arr)[(size_t)(n) * 3 + (d) - 1])
    for (int i = 0; i < N; i++) {
        double K[3] = {GIJ(c, i, 1, 1) / c->prm.rho, GIJ(c, i, 2, 2) / c->prm.rho, GIJ(c, i, 3, 3) / c->prm.rho};
        for (int d = 1; d <= 3; d++) C3(CC, d, i) = -(1. / (3. * K[d - 1]));
        Det[i] = 1. / (C3(CC, 1, i) * C3(CC, 2, i) + C3(CC, 2, i) * C3(CC, 3, i) + C3(CC, 1, i) * C3(CC, 3, i));
        for (int d = 1; d <= 3; d++) {
            C3(EE, d, i) = Det[i] / 2.;
            C3(FF, d, i) = Det[i] / 2.;
            C3(HH, d, i) = 1. / K[d - 1];
            C3(JJ, d, i) = -(1. / K[d - 1]);
        }
    }
    for (int i = 0; i < N; i++) {
        int i2;
        if ((i2 = NB(c, i, 2)) >= 0) {
            double gp = GF(c, i, 1, 1, 2), gm = GF(c, i2, 1, 1, 1);
            double *A = c->AA_px;
            A15(A, i, 1) = -(GG + (C3(CC, 2, i) + C3(CC, 3, i)) * C3(EE, 1, i) * C3(HH, 1, i)) * gp;
            A15(A, i, 2) = (GG - (C3(CC, 2, i) + C3(CC, 3, i)) * C3(FF, 1, i) * C3(HH, 1, i)) * gp
                           + (II + (C3(CC, 2, i2) + C3(CC, 3, i2)) * C3(EE, 1, i2) * C3(JJ, 1, i2)) * gm;
            A15(A, i, 3) = -(II - (C3(CC, 2, i2) + C3(CC, 3, i2)) * C3(FF, 1, i2) * C3(JJ, 1, i2)) * gm;
            A15(A, i, 4) = C3(CC, 3, i) * C3(EE, 2, i) * C3(HH, 1, i) * gp;
            A15(A, i, 5) = C3(CC, 3, i) * C3(FF, 2, i) * C3(HH, 1, i) * gp;
            A15(A, i, 6) = -(C3(CC, 3, i2) * C3(EE, 2, i2) * C3(JJ, 1, i2) * gm);
            A15(A, i, 7) = -(C3(CC, 3, i2) * C3(FF, 2, i2) * C3(JJ, 1, i2) * gm);
            A15(A, i, 8) = C3(CC, 2, i) * C3(EE, 3, i) * C3(HH, 1, i) * gp;
            A15(A, i, 9) = C3(CC, 2, i) * C3(FF, 3, i) * C3(HH, 1, i) * gp;
            A15(A, i, 10) = -(C3(CC, 2, i2) * C3(EE, 3, i2) * C3(JJ, 1, i2) * gm);
            A15(A, i, 11) = -(C3(CC, 2, i2) * C3(FF, 3, i2) * C3(JJ, 1, i2) * gm);
            A15(A, i, 12) = C3(CC, 2, i) * C3(CC, 3, i) * Det[i] * C3(HH, 1, i) * gp;
            A15(A, i, 13) = -(C3(CC, 2, i) * C3(CC, 3, i) * Det[i] * C3(HH, 1, i) * gp);
            A15(A, i, 14) = -(C3(CC, 2, i2) * C3(CC, 3, i2) * Det[i2] * C3(JJ, 1, i2) * gm);
            A15(A, i, 15) = C3(CC, 2, i2) * C3(CC, 3, i2) * Det[i2] * C3(JJ, 1, i2) * gm;
        }
        if ((i2 = NB(c, i, 4)) >= 0) {
            double gp = GF(c, i, 2, 2, 4), gm = GF(c, i2, 2, 2, 3);
            double *A = c->AA_py;
            A15(A, i, 1) = C3(CC, 3, i) * C3(EE, 1, i) * C3(HH, 2, i) * gp;
            A15(A, i, 2) = C3(CC, 3, i) * C3(FF, 1, i) * C3(HH, 2, i) * gp;
            A15(A, i, 3) = -(C3(CC, 3, i2) * C3(EE, 1, i2) * C3(JJ, 2, i2) * gm);
            A15(A, i, 4) = -(C3(CC, 3, i2) * C3(FF, 1, i2) * C3(JJ, 2, i2) * gm);
            A15(A, i, 5) = -(GG + (C3(CC, 1, i) + C3(CC, 3, i)) * C3(EE, 2, i) * C3(HH, 2, i)) * gp;
            A15(A, i, 6) = (GG - (C3(CC, 1, i) + C3(CC, 3, i)) * C3(FF, 2, i) * C3(HH, 2, i)) * gp
                           + (II + (C3(CC, 1, i2) + C3(CC, 3, i2)) * C3(EE, 2, i2) * C3(JJ, 2, i2)) * gm;
            A15(A, i, 7) = -(II - (C3(CC, 1, i2) + C3(CC, 3, i2)) * C3(FF, 2, i2) * C3(JJ, 2, i2)) * gm;
            A15(A, i, 8) = C3(CC, 1, i) * C3(EE, 3, i) * C3(HH, 2, i) * gp;
            A15(A, i, 9) = C3(CC, 1, i) * C3(FF, 3, i) * C3(HH, 2, i) * gp;
            A15(A, i, 10) = -(C3(CC, 1, i2) * C3(EE, 3, i2) * C3(JJ, 2, i2) * gm);
            A15(A, i, 11) = -(C3(CC, 1, i2) * C3(FF, 3, i2) * C3(JJ, 2, i2) * gm);
            A15(A, i, 12) = C3(CC, 1, i) * C3(CC, 3, i) * Det[i] * C3(HH, 2, i) * gp;
            A15(A, i, 13) = -(C3(CC, 1, i) * C3(CC, 3, i) * Det[i] * C3(HH, 2, i) * gp);
            A15(A, i, 14) = -(C3(CC, 1, i2) * C3(CC, 3, i2) * Det[i2] * C3(JJ, 2, i2) * gm);
            A15(A, i, 15) = C3(CC, 1, i2) * C3(CC, 3, i2) * Det[i2] * C3(JJ, 2, i2) * gm;
        }
        if ((i2 = NB(c, i, 6)) >= 0) {
            double gp = GF(c, i, 3, 3, 6), gm = GF(c, i2, 3, 3, 5);
            double *A = c->AA_pz;
            A15(A, i, 1) = C3(CC, 2, i) * C3(EE, 1, i) * C3(HH, 3, i) * gp;
            A15(A, i, 2) = C3(CC, 2, i) * C3(FF, 1, i) * C3(HH, 3, i) * gp;
            A15(A, i, 3) = -(C3(CC, 2, i2) * C3(EE, 1, i2) * C3(JJ, 3, i2) * gm);
            A15(A, i, 4) = -(C3(CC, 2, i2) * C3(FF, 1, i2) * C3(JJ, 3, i2) * gm);
            A15(A, i, 5) = C3(CC, 1, i) * C3(EE, 2, i) * C3(HH, 3, i) * gp;
            A15(A, i, 6) = C3(CC, 1, i) * C3(FF, 2, i) * C3(HH, 3, i) * gp;
            A15(A, i, 7) = -(C3(CC, 1, i2) * C3(EE, 2, i2) * C3(JJ, 3, i2) * gm);
            A15(A, i, 8) = -(C3(CC, 1, i2) * C3(FF, 2, i2) * C3(JJ, 3, i2) * gm);
            A15(A, i, 9) = -(GG + (C3(CC, 1, i) + C3(CC, 2, i)) * C3(EE, 3, i) * C3(HH, 3, i)) * gp;
            A15(A, i, 10) = (GG - (C3(CC, 1, i) + C3(CC, 2, i)) * C3(FF, 3, i) * C3(HH, 3, i)) * gp
                            + (II + (C3(CC, 1, i2) + C3(CC, 2, i2)) * C3(EE, 3, i2) * C3(JJ, 3, i2)) * gm;
            A15(A, i, 11) = -(II - (C3(CC, 1, i2) + C3(CC, 2, i2)) * C3(FF, 3, i2) * C3(JJ, 3, i2)) * gm;
            A15(A, i, 12) = C3(CC, 1, i) * C3(CC, 2, i) * Det[i] * C3(HH, 3, i) * gp;
            A15(A, i, 13) = -(C3(CC, 1, i) * C3(CC, 2, i) * Det[i] * C3(HH, 3, i) * gp);
            A15(A, i, 14) = -(C3(CC, 1, i2) * C3(CC, 2, i2) * Det[i2] * C3(JJ, 3, i2) * gm);
            A15(A, i, 15) = C3(CC, 1, i2) * C3(CC, 2, i2) * Det[i2] * C3(JJ, 3, i2) * gm;
        }
    }
    /* wall-face PPE constants BCs_faces_ap (nim.f90:682-742) */
    for (int b = 0; b < c->nbc; b++) {
        int M = c->bc_mesh[b], pos = c->bc_pos[b];
        int d = (pos + 1) / 2; /* normal direction */
        int e1 = d == 1 ? 2 : 1, e2 = d == 3 ? 2 : 3; /* the two tangential directions */
        int plus = (pos % 2 == 0);
        double g = GF(c, M, d, d, pos);
        double L = plus ? C3(HH, d, M) : C3(JJ, d, M); /* H on the + face, J on the - face */
        double S0 = plus ? GG : II;
        double Ce = C3(CC, e1, M) + C3(CC, e2, M);
        /* coefficients of faces 2d-1 and 2d (the wall face and its opposite) */
        double a_self = -(S0 + Ce * C3(EE, d, M) * L) * g;
        double a_opp = (S0 - Ce * C3(FF, d, M) * L) * g;
        /* the two tangential pairs: C of the other tangential direction times E/F of this one */
        double t1m = C3(CC, e2, M) * C3(EE, e1, M) * L * g, t1p = C3(CC, e2, M) * C3(FF, e1, M) * L * g;
        double t2m = C3(CC, e1, M) * C3(EE, e2, M) * L * g, t2p = C3(CC, e1, M) * C3(FF, e2, M) * L * g;
        double s7 = C3(CC, e1, M) * C3(CC, e2, M) * Det[M] * L * g;
        double ap[9];
        /* slots follow the face numbering 1..6 of the cell: the normal pair
         * 2d-1, 2d, then the two tangential pairs */
        ap[2 * d - 1] = a_self;
        ap[2 * d] = a_opp;
        ap[2 * e1 - 1] = t1m; ap[2 * e1] = t1p;
        ap[2 * e2 - 1] = t2m; ap[2 * e2] = t2p;
        for (int m = 1; m <= 6; m++) BCAP(c, b, m) = ap[m];
        BCAP(c, b, 7) = s7;
        BCAP(c, b, 8) = -s7;
    }
#undef C3
    free(CC); free(EE); free(FF); free(HH); free(JJ); free(Det);
}

/* ------------------------------------------------------------------ */
/* sources                                                             */
/* ------------------------------------------------------------------ */
/* dphi_dx2 / dphi_dy2 / dphi_dz2 (nse.f90:2866-2903), component d */
static double dphi(const ora_ctx *c, const double *phi, int n, int d)
{
    double s = phi[0 + 1] * AREA(c, MF(c, n, 2)) * NRM(c, MF(c, n, 2), d)
               - phi[0] * AREA(c, MF(c, n, 1)) * NRM(c, MF(c, n, 1), d)
               + phi[3] * AREA(c, MF(c, n, 4)) * NRM(c, MF(c, n, 4), d)
               - phi[2] * AREA(c, MF(c, n, 3)) * NRM(c, MF(c, n, 3), d)
               + phi[5] * AREA(c, MF(c, n, 6)) * NRM(c, MF(c, n, 6), d)
               - phi[4] * AREA(c, MF(c, n, 5)) * NRM(c, MF(c, n, 5), d);
    return 1. / c->volume[n] * s;
}

static void gather6(const ora_ctx *c, const double *f, int n, double out[6])
{
    for (int k = 1; k <= 6; k++) out[k - 1] = f[MF(c, n, k)];
}

/* tangential derivative used by the cross-diffusion operator
 * (nse.f90:1643-1653 pattern): derivative along direction with neighbours
 * (n1 on the minus side, n2 on the plus side) of the face at position fpos */
static double tder(const ora_ctx *c, const double *f, int n, int fpos, int n1, int n2, int n3)
{
    if (n1 < 0) return (-3 * f[MF(c, n, fpos)] + 4 * f[MF(c, n2, fpos)] - f[MF(c, n3, fpos)]) / 4.;
    if (n2 < 0) return (3 * f[MF(c, n, fpos)] - 4 * f[MF(c, n1, fpos)] + f[MF(c, n3, fpos)]) / 4.;
    return (f[MF(c, n2, fpos)] - f[MF(c, n1, fpos)]) / 4.;
}

/* calculate_R_D2 / calculate_R_D_p2 / calculate_R_D_T (nse.f90:1628-2134):
 * coef*(G1 f_x + G2 f_y + G3 f_z - g12 f_xy - g21 f_yx - g13 f_xz - g31 f_zx
 * - g23 f_yz - g32 f_zy) with one-sided closures at walls */
static double cross_diffusion(const ora_ctx *c, const double *f, int n, double coef)
{
    int nj1 = NB(c, n, 3), nj2 = NB(c, n, 4), nj3 = -1;
    if (nj1 < 0) nj3 = NB(c, nj2, 4);
    if (nj2 < 0) nj3 = NB(c, nj1, 3);
    double y1 = tder(c, f, n, 1, nj1, nj2, nj3), y2 = tder(c, f, n, 2, nj1, nj2, nj3);
    double fxy = 0.5 * (y2 - y1);
    y1 = tder(c, f, n, 5, nj1, nj2, nj3); y2 = tder(c, f, n, 6, nj1, nj2, nj3);
    double fzy = 0.5 * (y2 - y1);
    int ni1 = NB(c, n, 1), ni2 = NB(c, n, 2), ni3 = -1;
    if (ni1 < 0) ni3 = NB(c, ni2, 2);
    if (ni2 < 0) ni3 = NB(c, ni1, 1);
    double x1 = tder(c, f, n, 3, ni1, ni2, ni3), x2 = tder(c, f, n, 4, ni1, ni2, ni3);
    double fyx = 0.5 * (x2 - x1);
    x1 = tder(c, f, n, 5, ni1, ni2, ni3); x2 = tder(c, f, n, 6, ni1, ni2, ni3);
    double fzx = 0.5 * (x2 - x1);
    int nk1 = NB(c, n, 5), nk2 = NB(c, n, 6), nk3 = -1;
    if (nk1 < 0) nk3 = NB(c, nk2, 6);
    if (nk2 < 0) nk3 = NB(c, nk1, 5);
    double z1 = tder(c, f, n, 1, nk1, nk2, nk3), z2 = tder(c, f, n, 2, nk1, nk2, nk3);
    double fxz = 0.5 * (z2 - z1);
    z1 = tder(c, f, n, 3, nk1, nk2, nk3); z2 = tder(c, f, n, 4, nk1, nk2, nk3);
    double fyz = 0.5 * (z2 - z1);
    double fx = (f[MF(c, n, 2)] - f[MF(c, n, 1)]) / 2.;
    double fy = (f[MF(c, n, 4)] - f[MF(c, n, 3)]) / 2.;
    double fz = (f[MF(c, n, 6)] - f[MF(c, n, 5)]) / 2.;
    return coef * (c->G1[n] * fx + c->G2[n] * fy + c->G3[n] * fz - GIJ(c, n, 1, 2) * fxy - GIJ(c, n, 2, 1) * fyx
                   - GIJ(c, n, 1, 3) * fxz - GIJ(c, n, 3, 1) * fzx - GIJ(c, n, 2, 3) * fyz - GIJ(c, n, 3, 2) * fzy);
}

/* update_velocity_sources (nse.f90:199-209): calculate_gforces (2484-2502),
 * calculate_R_b (1562-1592), calculate_R_c_new (1393-1434), calculate_R_D2 */
void ora_velocity_sources(ora_ctx *c)
{
    const double gb = c->prm.g_beta;
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int i = 0; i < c->nmesh; i++) c->meshes_bz[i] = -(gb * 0.5 * (c->Txyz[i] + c->Txyz_old[i]));
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int f = 0; f < c->nface; f++) c->faces_bz[f] = -(gb * c->faces_T[f]);
    const double rinv = 1 / c->prm.rho, Rinv = 1 / c->prm.Re;
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int n = 0; n < c->nmesh; n++) {
        double pf[6];
        gather6(c, c->faces_p, n, pf);
        c->R_B_u[n] = 0. + rinv * dphi(c, pf, n, 1);
        c->R_B_v[n] = 0. + rinv * dphi(c, pf, n, 2);
        c->R_B_w[n] = c->meshes_bz[n] + rinv * dphi(c, pf, n, 3);
        double du1 = (c->faces_u[MF(c, n, 2)] - c->faces_u[MF(c, n, 1)]) / 2.;
        double du2 = (c->faces_u[MF(c, n, 4)] - c->faces_u[MF(c, n, 3)]) / 2.;
        double du3 = (c->faces_u[MF(c, n, 6)] - c->faces_u[MF(c, n, 5)]) / 2.;
        double dv1 = (c->faces_v[MF(c, n, 2)] - c->faces_v[MF(c, n, 1)]) / 2.;
        double dv2 = (c->faces_v[MF(c, n, 4)] - c->faces_v[MF(c, n, 3)]) / 2.;
        double dv3 = (c->faces_v[MF(c, n, 6)] - c->faces_v[MF(c, n, 5)]) / 2.;
        double dw1 = (c->faces_w[MF(c, n, 2)] - c->faces_w[MF(c, n, 1)]) / 2.;
        double dw2 = (c->faces_w[MF(c, n, 4)] - c->faces_w[MF(c, n, 3)]) / 2.;
        double dw3 = (c->faces_w[MF(c, n, 6)] - c->faces_w[MF(c, n, 5)]) / 2.;
        double uc = c->uxyzt[n] * XI(c, n, 1, 1) + c->vxyzt[n] * XI(c, n, 1, 2) + c->wxyzt[n] * XI(c, n, 1, 3);
        double vc = c->uxyzt[n] * XI(c, n, 2, 1) + c->vxyzt[n] * XI(c, n, 2, 2) + c->wxyzt[n] * XI(c, n, 2, 3);
        double wc = c->uxyzt[n] * XI(c, n, 3, 1) + c->vxyzt[n] * XI(c, n, 3, 2) + c->wxyzt[n] * XI(c, n, 3, 3);
        double a = c->meshes_u[n] - uc, b = c->meshes_v[n] - vc, g = c->meshes_w[n] - wc;
        c->R_C_u[n] = -(a * du1 + b * du2 + g * du3);
        c->R_C_v[n] = -(a * dv1 + b * dv2 + g * dv3);
        c->R_C_w[n] = -(a * dw1 + b * dw2 + g * dw3);
        c->R_D_u[n] = cross_diffusion(c, c->faces_u, n, Rinv);
        c->R_D_v[n] = cross_diffusion(c, c->faces_v, n, Rinv);
        c->R_D_w[n] = cross_diffusion(c, c->faces_w, n, Rinv);
    }
}

/* update_pressure_sources (nse.f90:228-235): calculate_R_b_p (2362-2407) and
 * calculate_R_D_p2 (1863-1997). faces_bx = faces_by = 0 (nse.f90:2496-2497). */
void ora_pressure_sources(ora_ctx *c)
{
    const double zero6[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int n = 0; n < c->nmesh; n++) {
        double uf[6], vf[6], wf[6], bzf[6];
        gather6(c, c->faces_u, n, uf);
        gather6(c, c->faces_v, n, vf);
        gather6(c, c->faces_w, n, wf);
        gather6(c, c->faces_bz, n, bzf);
        double du_dx = dphi(c, uf, n, 1), du_dy = dphi(c, uf, n, 2), du_dz = dphi(c, uf, n, 3);
        double dv_dx = dphi(c, vf, n, 1), dv_dy = dphi(c, vf, n, 2), dv_dz = dphi(c, vf, n, 3);
        double dw_dx = dphi(c, wf, n, 1), dw_dy = dphi(c, wf, n, 2), dw_dz = dphi(c, wf, n, 3);
        double dbx_dx = dphi(c, zero6, n, 1), dby_dy = dphi(c, zero6, n, 2), dbz_dz = dphi(c, bzf, n, 3);
        double Dila = du_dx + dv_dy + dw_dz;
        c->dila[n] = Dila;
        c->R_B_p[n] = -(du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz) - 2.0 * (du_dy * dv_dx + dv_dz * dw_dy + dw_dx * du_dz)
                      - dbx_dx - dby_dy - dbz_dz + Dila / (2 * c->prm.DT);
        c->R_D_p[n] = cross_diffusion(c, c->faces_p, n, 1 / c->prm.rho);
    }
}

/* update_temperature_sources (nse.f90:249-256): R_b_T = q = 0 (is_q false,
 * parameters.f90:179) and calculate_R_D_T (2000-2134) */
void ora_temperature_sources(ora_ctx *c)
{
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
    for (int n = 0; n < c->nmesh; n++) c->R_D_T[n] = cross_diffusion(c, c->faces_T, n, c->prm.alpha);
}

/* ------------------------------------------------------------------ */
/* per-cell sweep units                                                */
/* ------------------------------------------------------------------ */
/* integrate_velocity_field_formulation2_i (nse.f90:831-1006), one field */
static void vel_cell(ora_ctx *c, int i, double *fc, const double *fold, double *ff, double Rsum, double w,
                     const double *F4)
{
    fc[i] = (1 - w) * fc[i] + (w / F8(F4, i, 1)) * (+F8(F4, i, 2) * fold[i] + F8(F4, i, 3) * ff[MF(c, i, 1)]
                                                     + F8(F4, i, 4) * ff[MF(c, i, 2)] + F8(F4, i, 5) * ff[MF(c, i, 3)]
                                                     + F8(F4, i, 6) * ff[MF(c, i, 4)] + F8(F4, i, 7) * ff[MF(c, i, 5)]
                                                     + F8(F4, i, 8) * ff[MF(c, i, 6)] + Rsum);
}

/* owned face `pos` (2, 4 or 6) of cell i for one transported field; same
 * formula for momentum (nse.f90:878-1005) and energy (nse.f90:1032-1088) */
static void transport_face(ora_ctx *c, int i, int pos, double *fc, const double *fold, double *ff, double w,
                           const double *Fd)
{
    int i2 = NB(c, i, pos);
    if (i2 < 0) return;
    double B3, B4;
    if (pos == 2) {
        B3 = (GF(c, i2, 1, 2, 1) - GF(c, i, 1, 2, 2)) * (ff[MF(c, i, 4)] - ff[MF(c, i, 3)] + ff[MF(c, i2, 4)] - ff[MF(c, i2, 3)]) / 4.0;
        B4 = (GF(c, i2, 1, 3, 1) - GF(c, i, 1, 3, 2)) * (ff[MF(c, i, 6)] - ff[MF(c, i, 5)] + ff[MF(c, i2, 6)] - ff[MF(c, i2, 5)]) / 4.0;
    } else if (pos == 4) {
        B3 = (GF(c, i2, 2, 1, 3) - GF(c, i, 2, 1, 4)) * (ff[MF(c, i, 2)] - ff[MF(c, i, 1)] + ff[MF(c, i2, 2)] - ff[MF(c, i2, 1)]) / 4.0;
        B4 = (GF(c, i2, 2, 3, 3) - GF(c, i, 2, 3, 4)) * (ff[MF(c, i, 6)] - ff[MF(c, i, 5)] + ff[MF(c, i2, 6)] - ff[MF(c, i2, 5)]) / 4.0;
    } else {
        B3 = (GF(c, i2, 3, 1, 5) - GF(c, i, 3, 1, 6)) * (ff[MF(c, i, 2)] - ff[MF(c, i, 1)] + ff[MF(c, i2, 2)] - ff[MF(c, i2, 1)]) / 4.0;
        B4 = (GF(c, i2, 3, 2, 5) - GF(c, i, 3, 2, 6)) * (ff[MF(c, i, 4)] - ff[MF(c, i, 3)] + ff[MF(c, i2, 4)] - ff[MF(c, i2, 3)]) / 4.0;
    }
    int own = MF(c, i, pos), opp = MF(c, i, pos - 1), far = MF(c, i2, pos);
    ff[own] = (1 - w) * ff[own] + (w / F5(Fd, i, 1)) * (+F5(Fd, i, 2) * ff[opp] + F5(Fd, i, 3) * ff[far]
                                                        + F5(Fd, i, 4) * (fc[i] + fold[i])
                                                        + F5(Fd, i, 5) * (fc[i2] + fold[i2]) - B3 - B4);
}

static void velocity_unit(ora_ctx *c, int i)
{
    const ora_params *p = &c->prm;
    vel_cell(c, i, c->uxyz, c->uxyz_old, c->faces_u, c->R_C_u[i] + c->R_D_u[i] + c->R_B_u[i], p->w_u, c->M_F4);
    vel_cell(c, i, c->vxyz, c->vxyz_old, c->faces_v, c->R_C_v[i] + c->R_D_v[i] + c->R_B_v[i], p->w_v, c->M_F4);
    vel_cell(c, i, c->wxyz, c->wxyz_old, c->faces_w, c->R_C_w[i] + c->R_D_w[i] + c->R_B_w[i], p->w_w, c->M_F4);
    const int pos[3] = {2, 4, 6};
    double *Fd[3] = {c->M_F1, c->M_F2, c->M_F3};
    for (int d = 0; d < 3; d++) {
        transport_face(c, i, pos[d], c->uxyz, c->uxyz_old, c->faces_u, p->w_u, Fd[d]);
        transport_face(c, i, pos[d], c->vxyz, c->vxyz_old, c->faces_v, p->w_v, Fd[d]);
        transport_face(c, i, pos[d], c->wxyz, c->wxyz_old, c->faces_w, p->w_w, Fd[d]);
    }
}

static void temperature_unit(ora_ctx *c, int i)
{
    double w = c->prm.w_T;
    vel_cell(c, i, c->Txyz, c->Txyz_old, c->faces_T, c->R_D_T[i] + 0., w, c->MT_F4);
    transport_face(c, i, 2, c->Txyz, c->Txyz_old, c->faces_T, w, c->MT_F1);
    transport_face(c, i, 4, c->Txyz, c->Txyz_old, c->faces_T, w, c->MT_F2);
    transport_face(c, i, 6, c->Txyz, c->Txyz_old, c->faces_T, w, c->MT_F3);
}

/* integrate_pressure_field_i (nse.f90:1091-1289): owned interior faces */
static void pressure_unit(ora_ctx *c, int n1)
{
    double *P = c->faces_p;
    const double w = c->prm.w_p;
    const double S1 = c->R_D_p[n1] + c->R_B_p[n1];
    int ni1 = NB(c, n1, 1), ni2 = NB(c, n1, 2), nj1 = NB(c, n1, 3), nj2 = NB(c, n1, 4);
    int nk1 = NB(c, n1, 5), nk2 = NB(c, n1, 6);
    int ni3 = -1, nj3 = -1, nk3 = -1;
    if (ni1 < 0) ni3 = NB(c, ni2, 2);
    if (ni2 < 0) ni3 = NB(c, ni1, 1);
    if (nj1 < 0) nj3 = NB(c, nj2, 4);
    if (nj2 < 0) nj3 = NB(c, nj1, 3);
    if (nk1 < 0) nk3 = NB(c, nk2, 6);
    if (nk2 < 0) nk3 = NB(c, nk1, 5);
#define PF(n, f) P[MF(c, n, f)]
    if (c->faces_bc[MF(c, n1, 2)] == 1) {
        int n2 = NB(c, n1, 2);
        const double *A = &c->AA_px[(size_t)n1 * 15] - 1;
        double B1 = S1 * A[13], B2 = (c->R_D_p[n2] + c->R_B_p[n2]) * A[15];
        double p_y = tder(c, P, n1, 2, nj1, nj2, nj3), p_z = tder(c, P, n1, 2, nk1, nk2, nk3);
        double B3 = (GF(c, n2, 1, 2, 1) - GF(c, n1, 1, 2, 2)) * p_y;
        double B4 = (GF(c, n2, 1, 3, 1) - GF(c, n1, 1, 3, 2)) * p_z;
        P[MF(c, n1, 2)] = (1 - w) * PF(n1, 2) + w / (A[2]) * (-A[1] * PF(n1, 1) - A[3] * PF(n2, 2) - A[4] * PF(n1, 3)
                                                               - A[5] * PF(n1, 4) - A[6] * PF(n2, 3) - A[7] * PF(n2, 4)
                                                               - A[8] * PF(n1, 5) - A[9] * PF(n1, 6) - A[10] * PF(n2, 5)
                                                               - A[11] * PF(n2, 6) + B1 + B2 + B3 + B4);
    }
    if (c->faces_bc[MF(c, n1, 4)] == 1) {
        int n2 = NB(c, n1, 4);
        const double *A = &c->AA_py[(size_t)n1 * 15] - 1;
        double B1 = S1 * A[13], B2 = (c->R_D_p[n2] + c->R_B_p[n2]) * A[15];
        double p_x = tder(c, P, n1, 4, ni1, ni2, ni3), p_z = tder(c, P, n1, 4, nk1, nk2, nk3);
        double B3 = (GF(c, n2, 2, 1, 3) - GF(c, n1, 2, 1, 4)) * p_x;
        double B4 = (GF(c, n2, 2, 3, 3) - GF(c, n1, 2, 3, 4)) * p_z;
        P[MF(c, n1, 4)] = (1 - w) * PF(n1, 4) + w / (A[6]) * (-A[1] * PF(n1, 1) - A[2] * PF(n1, 2) - A[3] * PF(n2, 1)
                                                               - A[4] * PF(n2, 2) - A[5] * PF(n1, 3) - A[7] * PF(n2, 4)
                                                               - A[8] * PF(n1, 5) - A[9] * PF(n1, 6) - A[10] * PF(n2, 5)
                                                               - A[11] * PF(n2, 6) + B1 + B2 + B3 + B4);
    }
    if (c->faces_bc[MF(c, n1, 6)] == 1) {
        int n2 = NB(c, n1, 6);
        const double *A = &c->AA_pz[(size_t)n1 * 15] - 1;
        double B1 = S1 * A[13], B2 = (c->R_D_p[n2] + c->R_B_p[n2]) * A[15];
        double p_x = tder(c, P, n1, 6, ni1, ni2, ni3), p_y = tder(c, P, n1, 6, nj1, nj2, nj3);
        double B3 = (GF(c, n2, 3, 1, 5) - GF(c, n1, 3, 1, 6)) * p_x;
        double B4 = (GF(c, n2, 3, 2, 5) - GF(c, n1, 3, 2, 6)) * p_y;
        P[MF(c, n1, 6)] = (1 - w) * PF(n1, 6) + w / (A[10]) * (-A[1] * PF(n1, 1) - A[2] * PF(n1, 2) - A[3] * PF(n2, 1)
                                                                - A[4] * PF(n2, 2) - A[5] * PF(n1, 3) - A[6] * PF(n1, 4)
                                                                - A[7] * PF(n2, 3) - A[8] * PF(n2, 4) - A[9] * PF(n1, 5)
                                                                - A[11] * PF(n2, 6) + B1 + B2 + B3 + B4);
    }
#undef PF
}

/* apply_pressure_bc for one boundary face b (bc.f90:147-341) */
static void pressure_wall(ora_ctx *c, int b)
{
    double *P = c->faces_p;
    const double nu = c->prm.nu, rho = c->prm.rho, w = c->prm.w_p2;
    int M1 = c->bc_mesh[b], fn = c->bc_pos[b], f = c->bc_faces[b];
    double n1 = NRM(c, f, 1), n2 = NRM(c, f, 2), n3 = NRM(c, f, 3);
    double b_n = 0. * n1 + 0. * n2 + c->faces_bz[f] * n3;
    double vn[5];
    for (int j = 1; j <= 5; j++) {
        int q = BCPT(c, b, j);
        vn[j - 1] = c->faces_u[q] * n1 + c->faces_v[q] * n2 + c->faces_w[q] * n3;
    }
    /* the six prescribed-derivative terms (bc.f90:164-166) multiply the BC
     * functions BCs_derivatives_*, which are identically zero
     * (general.f90:289-420), and are omitted */
    double d2 = -(BCA(c, b, 1) * vn[0]) - BCA(c, b, 2) * vn[1] - BCA(c, b, 3) * vn[2] - BCA(c, b, 4) * vn[3]
                - BCA(c, b, 5) * vn[4];
    /* tangential derivative pairs: (neighbour slots, metric component) */
    int ta, tb; /* first and second tangential directions (neighbour slot pairs) */
    int ga1, gb1, ga2, gb2;
    if (fn <= 2) { ta = 3; tb = 5; ga1 = 1; gb1 = 2; ga2 = 1; gb2 = 3; }
    else if (fn <= 4) { ta = 1; tb = 5; ga1 = 2; gb1 = 1; ga2 = 2; gb2 = 3; }
    else { ta = 1; tb = 3; ga1 = 3; gb1 = 1; ga2 = 3; gb2 = 2; }
    double B2, B3;
    {
        int M2 = NB(c, M1, ta), M3 = NB(c, M1, ta + 1);
        double g = GF(c, M1, ga1, gb1, fn);
        if (M2 < 0)
            B2 = (-3. * P[MF(c, M1, fn)] + 4. * P[MF(c, M3, fn)] - P[MF(c, NB(c, M3, ta + 1), fn)]) / 4.0 * g;
        else if (M3 < 0)
            B2 = (3. * P[MF(c, M1, fn)] - 4. * P[MF(c, M2, fn)] + P[MF(c, NB(c, M2, ta), fn)]) / 4.0 * g;
        else {
            /* bc.f90:318: the top wall (position 6) uses gij_faces(3,2,...) in
             * its central branch; reproduced as in the reference */
            double gc = fn == 6 ? GF(c, M1, 3, 2, fn) : g;
            B2 = (P[MF(c, M3, fn)] - P[MF(c, M2, fn)]) / 4.0 * gc;
        }
    }
    {
        int M2 = NB(c, M1, tb), M3 = NB(c, M1, tb + 1);
        double g = GF(c, M1, ga2, gb2, fn);
        if (M2 < 0)
            B3 = (-3. * P[MF(c, M1, fn)] + 4. * P[MF(c, M3, fn)] - P[MF(c, NB(c, M3, tb + 1), fn)]) / 4.0 * g;
        else if (M3 < 0)
            B3 = (3. * P[MF(c, M1, fn)] - 4. * P[MF(c, M2, fn)] + P[MF(c, NB(c, M2, tb), fn)]) / 4.0 * g;
        else
            B3 = (P[MF(c, M3, fn)] - P[MF(c, M2, fn)]) / 4.0 * g;
    }
    double B1 = BCAP(c, b, 8) * (c->R_D_p[M1] + c->R_B_p[M1]);
    double s = 0.;
    int first = 1;
    for (int k = 1; k <= 6; k++) {
        if (k == fn) continue;
        double t = BCAP(c, b, k) * P[MF(c, M1, k)];
        s = first ? -t : s - t;
        first = 0;
    }
    int own = MF(c, M1, fn);
    P[own] = (1 - w) * P[own] + w / (BCAP(c, b, fn)) * (s + B1 - B2 - B3 + (nu * rho * d2 - b_n * rho));
}

void ora_pressure_bc(ora_ctx *c)
{
    for (int b = 0; b < c->nbc; b++) pressure_wall(c, b);
}

/* apply_zero_flux (bc.f90:345-390): adiabatic top and bottom walls */
static double hdist(const ora_ctx *c, int f1, int f2)
{
    double dx = MID(c, f1, 1) - MID(c, f2, 1), dy = MID(c, f1, 2) - MID(c, f2, 2), dz = MID(c, f1, 3) - MID(c, f2, 3);
    return sqrt(dx * dx + dy * dy + dz * dz);
}

void ora_zero_flux(ora_ctx *c)
{
    double *T = c->faces_T;
    for (int b = 0; b < c->nbc; b++) {
        int pos = c->bc_pos[b];
        if (pos != 5 && pos != 6) continue;
        int M = c->bc_mesh[b], s = pos == 5 ? 6 : 5;
        int f1 = c->bc_faces[b], f2 = MF(c, M, s), f3 = MF(c, NB(c, M, s), s), f4 = MF(c, NB(c, NB(c, M, s), s), s);
        double h1 = hdist(c, f1, f2), h2 = hdist(c, f2, f3), h3 = hdist(c, f3, f4);
        T[f1] = (+((h1 + h2) * (h1 + h2 + h3)) / (h1 * h2 * (h2 + h3)) * T[f2]
                 - (h1 * (h1 + h2 + h3)) / (h2 * h3 * (h1 + h2)) * T[f3]
                 + (h1 * (h1 + h2)) / (h3 * (h2 + h3) * (h1 + h2 + h3)) * T[f4])
                / (1. / h1 + 1. / (h1 + h2) + 1. / (h1 + h2 + h3));
    }
}

/* ------------------------------------------------------------------ */
/* sweeps in either order                                              */
/* ------------------------------------------------------------------ */
void ora_velocity_sweep(ora_ctx *c, int order)
{
    if (order == ORA_NATURAL) {
        for (int i = 0; i < c->nmesh; i++) velocity_unit(c, i);
    } else {
        for (int col = 0; col < 4; col++) {
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
            for (int i = 0; i < c->nmesh; i++)
                if (cell_color(c, i) == col) velocity_unit(c, i);
        }
    }
}

void ora_temperature_sweep(ora_ctx *c, int order)
{
    if (order == ORA_NATURAL) {
        for (int i = 0; i < c->nmesh; i++) temperature_unit(c, i);
    } else {
        for (int col = 0; col < 4; col++) {
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
            for (int i = 0; i < c->nmesh; i++)
                if (cell_color(c, i) == col) temperature_unit(c, i);
        }
    }
}

/* natural: apply_pressure_bc then integrate_pressure_field (nse.f90:
 * 229-231); colored: per cell its wall faces (boundary group order) then
 * its interior faces, colour by colour */
void ora_pressure_stage(ora_ctx *c, int order)
{
    if (order == ORA_NATURAL) {
        ora_pressure_bc(c);
        for (int i = 0; i < c->nmesh; i++) pressure_unit(c, i);
        return;
    }
    /* wall faces of each cell in boundary-group order: bottom (5), top (6),
     * hexagon (1), cylinder (2) */
    int *first = (int *)malloc(sizeof(int) * c->nmesh * 4);
    for (int i = 0; i < c->nmesh * 4; i++) first[i] = -1;
    for (int b = 0; b < c->nbc; b++) {
        int slot = c->bc_pos[b] == 5 ? 0 : c->bc_pos[b] == 6 ? 1 : c->bc_pos[b] == 1 ? 2 : 3;
        first[c->bc_mesh[b] * 4 + slot] = b;
    }
    for (int col = 0; col < 4; col++) {
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN)
        for (int i = 0; i < c->nmesh; i++) {
            if (cell_color(c, i) != col) continue;
            for (int s = 0; s < 4; s++)
                if (first[i * 4 + s] >= 0) pressure_wall(c, first[i * 4 + s]);
            pressure_unit(c, i);
        }
    }
    free(first);
}

/* integrate_pressure_field alone (nse.f90:216-226), natural order */
void ora_pressure_sweep(ora_ctx *c)
{
    for (int i = 0; i < c->nmesh; i++) pressure_unit(c, i);
}

void ora_set_p_measure(ora_ctx *c, int mode) { c->p_measure = mode; }

/* find_error (nse.f90:260-325); with p_measure 1 the pressure increment is
 * taken as half its range over the faces instead of its largest magnitude */
void ora_find_error(ora_ctx *c, double err[5])
{
    double eu = 0, ev = 0, ew = 0, ep = 0, eT = 0, pmax = -INFINITY, pmin = INFINITY;
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN) reduction(max : eu, ev, ew, ep, eT, pmax) reduction(min : pmin)
    for (int i = 0; i < c->nface; i++) {
        const double dp = c->faces_p[i] - c->faces_p_old[i];
        eu = fmax(eu, fabs(c->faces_u[i] - c->faces_u_old[i]));
        ev = fmax(ev, fabs(c->faces_v[i] - c->faces_v_old[i]));
        ew = fmax(ew, fabs(c->faces_w[i] - c->faces_w_old[i]));
        ep = fmax(ep, fabs(dp));
        pmax = fmax(pmax, dp);
        pmin = fmin(pmin, dp);
        eT = fmax(eT, fabs(c->faces_T[i] - c->faces_T_old[i]));
        c->faces_u_old[i] = c->faces_u[i];
        c->faces_v_old[i] = c->faces_v[i];
        c->faces_w_old[i] = c->faces_w[i];
        c->faces_p_old[i] = c->faces_p[i];
        c->faces_T_old[i] = c->faces_T[i];
    }
    double cu = 0, cv = 0, cw = 0, cT = 0;
#pragma omp parallel for schedule(static) if (c->nmesh >= ORA_PAR_MIN) reduction(max : cu, cv, cw, cT)
    for (int i = 0; i < c->nmesh; i++) {
        cu = fmax(cu, fabs(c->uxyz[i] - c->uxyz_old_con[i]));
        cv = fmax(cv, fabs(c->vxyz[i] - c->vxyz_old_con[i]));
        cw = fmax(cw, fabs(c->wxyz[i] - c->wxyz_old_con[i]));
        cT = fmax(cT, fabs(c->Txyz[i] - c->Txyz_old_con[i]));
        c->uxyz_old_con[i] = c->uxyz[i];
        c->vxyz_old_con[i] = c->vxyz[i];
        c->wxyz_old_con[i] = c->wxyz[i];
        c->Txyz_old_con[i] = c->Txyz[i];
    }
    err[0] = fmax(eu, cu); err[1] = fmax(ev, cv); err[2] = fmax(ew, cw); err[3] = ep; err[4] = fmax(eT, cT);
    if (c->p_measure == 1) err[3] = isfinite(ep) ? (pmax - pmin) / 2 : ep;
}

static void snapshot_old(ora_ctx *c)
{
    memcpy(c->faces_u_old, c->faces_u, sizeof(double) * c->nface);
    memcpy(c->faces_v_old, c->faces_v, sizeof(double) * c->nface);
    memcpy(c->faces_w_old, c->faces_w, sizeof(double) * c->nface);
    memcpy(c->faces_p_old, c->faces_p, sizeof(double) * c->nface);
    memcpy(c->faces_T_old, c->faces_T, sizeof(double) * c->nface);
    memcpy(c->uxyz_old_con, c->uxyz, sizeof(double) * c->nmesh);
    memcpy(c->vxyz_old_con, c->vxyz, sizeof(double) * c->nmesh);
    memcpy(c->wxyz_old_con, c->wxyz, sizeof(double) * c->nmesh);
    memcpy(c->Txyz_old_con, c->Txyz, sizeof(double) * c->nmesh);
}

/* one outer iteration of integrate_nse (nse.f90:134-162) */
void ora_iteration(ora_ctx *c, int order, double err[5])
{
    ora_velocity_sources(c);
    ora_velocity_sweep(c, order);
    ora_pressure_sources(c);
    ora_pressure_stage(c, order);
    ora_temperature_constants(c);
    ora_zero_flux(c);
    ora_temperature_sources(c);
    ora_temperature_sweep(c, order);
    ora_find_error(c, err);
}

/* check_steady_state (nse.f90:1316-1332) */
void ora_check_steady_state(ora_ctx *c)
{
    double du = 0, dv = 0, dw = 0, dT = 0, den = 2 * c->prm.DT;
    for (int i = 0; i < c->nmesh; i++) {
        du = fmax(du, fabs(c->uxyz[i] - c->uxyz_old[i]) / den);
        dv = fmax(dv, fabs(c->vxyz[i] - c->vxyz_old[i]) / den);
        dw = fmax(dw, fabs(c->wxyz[i] - c->wxyz_old[i]) / den);
        dT = fmax(dT, fabs(c->Txyz[i] - c->Txyz_old[i]) / den);
    }
    c->dudt_max = fmax(dw, fmax(du, dv));
    c->dTdt_max = dT;
}

/* evaluate_D (nse.f90:2504-2537): RMS dilatation */
void ora_evaluate_D(ora_ctx *c)
{
    double sum1 = 0.0;
    for (int n = 0; n < c->nmesh; n++) {
        double uf[6], vf[6], wf[6];
        gather6(c, c->faces_u, n, uf);
        gather6(c, c->faces_v, n, vf);
        gather6(c, c->faces_w, n, wf);
        double D = dphi(c, uf, n, 1) + dphi(c, vf, n, 2) + dphi(c, wf, n, 3);
        c->dila[n] = D;
        sum1 = sum1 + D * D;
    }
    c->Dilatation = sqrt(sum1 / c->nmesh);
}

/* integrate_nse (nse.f90:117-176). Returns 0, or 1 if max_iter was reached
 * (the reference has no cap), or 2 on a non-finite error. In colored mode the
 * change is measured from the state at entry and at least min_iter passes run. */
int ora_inner_solve(ora_ctx *c, int order, const double tol[5], int max_iter, int min_iter, int *iters,
                    double err[5])
{
    int cnt = 0;
    for (int k = 0; k < 5; k++) err[k] = 1.;
    if (order == ORA_COLORED) snapshot_old(c);
    while (err[0] > tol[0] || err[1] > tol[1] || err[2] > tol[2] || err[3] > tol[3] || err[4] > tol[4]
           || cnt < min_iter) {
        if (cnt >= max_iter) { *iters = cnt; return 1; }
        cnt++;
        ora_iteration(c, order, err);
        for (int k = 0; k < 5; k++)
            if (!isfinite(err[k])) { *iters = cnt; return 2; }
    }
    ora_check_steady_state(c);
    ora_evaluate_D(c);
    *iters = cnt;
    return 0;
}

/* update_time_step (nse.f90:1292-1313) */
void ora_update_time_step(ora_ctx *c)
{
    c->STEP = c->STEP + 1;
    c->TI = c->TI + 2 * c->prm.DT;
    for (int i = 0; i < c->nmesh; i++) {
        c->uxyzt[i] = (c->uxyz_old[i] + c->uxyz[i]) / 2.;
        c->uxyz_old[i] = c->uxyz[i];
        c->vxyzt[i] = (c->vxyz_old[i] + c->vxyz[i]) / 2.;
        c->vxyz_old[i] = c->vxyz[i];
        c->wxyzt[i] = (c->wxyz_old[i] + c->wxyz[i]) / 2.;
        c->wxyz_old[i] = c->wxyz[i];
        c->Txyzt[i] = (c->Txyz_old[i] + c->Txyz[i]) / 2.;
        c->Txyz_old[i] = c->Txyz[i];
        c->meshes_p[i] = (c->faces_p[MF(c, i, 1)] + c->faces_p[MF(c, i, 2)] + c->faces_p[MF(c, i, 3)]
                          + c->faces_p[MF(c, i, 4)] + c->faces_p[MF(c, i, 5)] + c->faces_p[MF(c, i, 6)]) / 6.0;
        c->dila_old[i] = c->dila[i];
    }
}

/* find_Nu2 (nse.f90:2699-2798). out = {Nu_H, Nu_C, Average, Area1, Area2}
 * as printed by the reference (nse.f90:2792-2796): index 0 is the outer
 * cylinder (position 2) scaled by R_o, index 1 the hexagon (position 1)
 * scaled by R_i. */
void ora_nusselt(ora_ctx *c, double out[5])
{
    double sum1 = 0., sum2 = 0., sumA1 = 0., sumA2 = 0.;
    const double *T = c->faces_T;
    for (int b = 0; b < c->nbc; b++) {
        int pos = c->bc_pos[b];
        if (pos != 1 && pos != 2) continue;
        int M = c->bc_mesh[b], s = pos == 1 ? 2 : 1;
        int f1 = c->bc_faces[b], f2 = MF(c, M, s), f3 = MF(c, NB(c, M, s), s), f4 = MF(c, NB(c, NB(c, M, s), s), s);
        double h1 = hdist(c, f1, f2), h2 = hdist(c, f2, f3), h3 = hdist(c, f3, f4);
        double d = (-(1. / h1 + 1. / (h1 + h2) + 1. / (h1 + h2 + h3)) * T[f1]
                    + ((h1 + h2) * (h1 + h2 + h3)) / (h1 * h2 * (h2 + h3)) * T[f2]
                    - (h1 * (h1 + h2 + h3)) / (h2 * h3 * (h1 + h2)) * T[f3]
                    + (h1 * (h1 + h2)) / (h3 * (h2 + h3) * (h1 + h2 + h3)) * T[f4]);
        if (pos == 1) {
            double dTdx = -d;
            sum2 = sum2 + AREA(c, MF(c, M, 1)) * dTdx;
            sumA2 = sumA2 + AREA(c, MF(c, M, 1));
        } else {
            sum1 = sum1 + AREA(c, MF(c, M, 2)) * d;
            sumA1 = sumA1 + AREA(c, MF(c, M, 2));
        }
    }
    out[0] = -(sum1 / sumA1 * c->prm.R_o);
    out[1] = -(sum2 / sumA2 * c->prm.R_i);
    out[2] = -((sum1 / sumA1 * c->prm.R_o) + (sum2 / sumA2 * c->prm.R_i)) / 2.;
    out[3] = sumA1;
    out[4] = sumA2;
}

/* one time step of solve_nse (nse.f90:49-73) without the I/O; returns the
 * inner-solve status. scal = {iters, err[5], Dilatation, dudt, dTdt}. */
int ora_time_step(ora_ctx *c, int order, const double tol[5], int max_iter, int min_iter, double scal[9],
                  double nu[5])
{
    int iters = 0;
    double err[5];
    ora_velocity_constants(c);
    int rc = ora_inner_solve(c, order, tol, max_iter, min_iter, &iters, err);
    scal[0] = iters;
    for (int k = 0; k < 5; k++) scal[1 + k] = err[k];
    scal[6] = c->Dilatation; scal[7] = c->dudt_max; scal[8] = c->dTdt_max;
    if (rc) return rc;
    ora_update_time_step(c);
    ora_nusselt(c, nu);
    return 0;
}
