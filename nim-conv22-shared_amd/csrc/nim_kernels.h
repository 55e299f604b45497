// Device-side data layout and helpers of the NIM time-step kernels (gfx950).
//
// Layout in HBM (structure of arrays, FP64):
//  * cells: c = P(i) + NT*(j + NR*k), i azimuthal (periodic), j radial, k
//    axial, P(i) = (i mod 2)*NT/2 + i/2: the reference's natural order
//    (create_geometry.f90:449-473) with each azimuthal row split into its
//    even and odd cells, so that every colour phase of a sweep reads and
//    writes contiguous half-rows.
//  * faces of one field: three families in one buffer of NF doubles
//      radial  R(i,j,k), j in [0,NR]   at P(i) + NT*(j + (NR+1)*k)
//      azimuthal T(i,j,k)              at oT + P(i) + NT*(j + NR*k)
//      axial   Z(i,j,k), k in [0,NZ]   at oZ + P(i) + NT*(j + NR*k)
//    Cell-local face p (reference frame, mesh.f90:65-69): 1 = R(i,j), 2 =
//    R(i,j+1), 3 = T(i), 4 = T(i+1 mod NT), 5 = Z(k), 6 = Z(k+1).
//  * per-cell arrays grouped as [slot][N] blocks (one allocation per group).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

namespace nim {

// per-cell geometry slots in Dev::gc
enum {
    G_INVV = 0,
    G_XI = 1,     // xi_x(a,b) at G_XI + 3*(a-1) + (b-1)
    G_GD = 10,    // g11, g22, g33
    G_GO = 13,    // g12, g21, g13, g31, g23, g32
    G_G = 19,     // G1, G2, G3
    G_GF = 22,    // gij_faces(1,1,1), (1,1,2), (2,2,3), (2,2,4), (3,3,5), (3,3,6)
    G_DIFF = 28,  // d2a, d2b, d4a, d4b, d6a, d6b (neighbour-minus-own face metric differences)
    NGC = 34
};

// per-wall-face slots in Dev::wall (boundary faces in the reference's BC group
// order: bottom, top, hexagon, cylinder)
enum {
    W_A = 0,     // BCs_faces_a(1..5)
    W_AP = 5,    // BCs_faces_ap(1..8)
    W_N = 13,    // face normal
    W_GB2 = 16,  // metric factor of the first tangential derivative, one-sided branches
    W_GB2C = 17, // same, central branch (differs on the top wall, bc.f90:318)
    W_GB3 = 18,  // metric factor of the second tangential derivative
    W_ZF = 19,   // zero-flux extrapolation: c2, c3, c4, den
    W_NU = 23,   // Nusselt one-sided derivative: S, A, B, C, wall area
    NWS = 28
};

// per-cell state slots in Dev::sc
enum {
    S_CUR = 0,   // u, v, w, T
    S_OLD = 4,   // previous time level
    S_TC = 8,    // time-centred
    S_P = 12,
    S_DILA = 13,
    S_DILAO = 14,
    NSC = 15
};

// coefficient slots in Dev::MF / Dev::TF: F4(1..8), F1(1..5), F2(1..5), F3(1..5)
enum { C_F4 = 0, C_F1 = 8, C_F2 = 13, C_F3 = 18, NCF = 23 };
// pressure coefficients read by the sweep, per direction in Dev::AQ
// ([3][NAQ][N]): 9 of AA_p? (15 slots, nim.f90:622-697). At u = 0 the
// coefficients E and F are equal, so each tangential pair of slots holds one
// value twice and is stored once; AA(12) and AA(14) are not read.
//   SELF, DIAG, FAR   the normal pair and the diagonal of the face equation
//   TA0, TB0          first tangential direction, own and partner side
//   TA1, TB1          second tangential direction, own and partner side
//   S1, S2            AA(13) and AA(15), the weights of the two cell sources
enum { Q_SELF = 0, Q_DIAG = 1, Q_FAR = 2, Q_TA0 = 3, Q_TB0 = 4, Q_TA1 = 5, Q_TB1 = 6, Q_S1 = 7, Q_S2 = 8, NAQ = 9 };
// source slots in Dev::src: R sums for u, v, w; pressure R_D_p + R_b_p; R_D_T
enum { R_U = 0, R_V = 1, R_W = 2, R_P = 3, R_T = 4, NSRC = 5 };
// face field slots in Dev::fc
enum { F_U = 0, F_V = 1, F_W = 2, F_P = 3, F_T = 4 };

// Load of data that a sweep reads once per pass (coefficients, metric
// differences, FP32 residuals), marked non-temporal when NT is set
// (global_load ... nt): the caches then favour the faces and cells that the
// other colour phases and the next kernels read again. The value is the same
// either way. NT is a template argument: a run-time choice between the two
// loads of one address can be merged into a plain load before it is folded.
template <bool NT = true, typename T>
__device__ __forceinline__ T ld_pass(const T *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

constexpr int ERR_COPIES = 64;  // spread of the per-iteration max-increment atomics
constexpr int ERR_STRIDE = 16;  // one 128-byte cache line per copy
// reduced per pass: max |increment| of u, v, w, p, T (fields 0-4), and the
// largest pressure increment and the largest negated one (5, 6; ordered
// keys of signed doubles), whose half sum is half the range of the
// pressure increments
enum { E_PHI = 5, E_PLO = 6, NERR = 7 };
__device__ __host__ inline size_t err_slot(int slot, int field, int copy)
{
    return ((size_t)(slot * NERR + field) * ERR_COPIES + copy) * ERR_STRIDE;
}

// pressure convergence measure (Ctl::p_measure):
//  P_MAXABS            max |dp| of the pass (or max |w r / a| of the FP64
//                      residual in the mixed loops)
//  P_HALFRANGE         half the range of the pass's increments dp, or of
//                      w r / a in the mixed loops
//  P_HALFRANGE_UPDATE  mixed loops: half the range of the last cycle's FP64
//                      update of p divided by its FP32 passes (the residual
//                      pass gives only max |w r / a|, used before the first
//                      update). The pressure level drifts by the same amount
//                      every pass, so the residual of the drifting rest state
//                      is that amount times the sweep's splitting applied to
//                      a constant, which has a range; the update's range is 0
enum { P_MAXABS = 0, P_HALFRANGE = 1, P_HALFRANGE_UPDATE = 2 };

struct Ctl {
    int32_t done, iter, status, max_iter, min_iter, p_measure, pad[2];
    double tol[5];
    double err[5];
};

struct Dev {
    int NT, NR, NZ, N, NF, NB, oT, oZ;
    double nu, Rinv, alpha, rinv, rho, g_beta, DT;
    double w_u, w_v, w_w, w_p, w_p2, w_T, v_limit;
    int kfast;          // per-cell kernels: blocks numbered plane-fastest (see cell_of_block)
    int rslab;          // per-cell kernels: one radial slab per XCD (see cell_of_block)
    int krev;           // with rslab: walk the planes from the top (see cell_of_block)
    const double *gc;   // [NGC][N]
    const double *An;   // [3][NF] face area times unit normal
    double *AA;         // [3][15][N] pressure coefficients AA_px, AA_py, AA_pz
    double *AQ;         // [3][NAQ][N] the ones the pressure sweep reads
    double *wall;       // [NWS][NB]
    const int32_t *wpts;// [5][NB] face slots of the 5-point normal stencil
    double *sc;         // [NSC][N]
    double *fc;         // [5][NF]
    double *MA, *TA;    // [18][N] exponential-fit constants (momentum, energy)
    double *MF, *TF;    // [NCF][N]
    double *src;        // [NSRC][N]
    double *red;        // reduction scratch
    Ctl *ctl;
    unsigned long long *errs;  // [2][NERR][ERR_COPIES] at err_slot()
};

// What a sweep routine does with the equation of one unknown, a phi = S(x):
//  GS     relax in place, phi <- (1 - w) phi + (w / a) S(x)   (FP64 pass)
//  RES    write the residual r = S(x0) - a phi0 without moving x0 (FP64)
//  DELTA  relax the correction, dphi <- (1 - w) dphi + (w / a)(S_lin(dx) + r)
//         where S_lin drops the terms that do not depend on the unknowns
//         (previous time level) and r is the residual of the last RES pass
//         (FP32). x0 + dx then follows the FP64 iteration's sequence, except
//         that the energy coefficients (exponential fits of the velocity)
//         stay at x0 until the next RES pass.
enum { MODE_GS = 0, MODE_RES = 1, MODE_DELTA = 2 };

// The arrays a kernel works on: the FP64 state and constants, or the FP32
// corrections with FP32 copies of the constants. Slot layouts are those of
// Dev (gc [NGC][N], AQ [3*NAQ][N], wall [NWS][NB], MF/TF [NCF][N], src
// [NSRC][N], cells [4][N] at S_CUR, faces [5][NF]).
template <typename T> struct View {
    const T *gc, *An, *AQ, *wall, *MF, *TF;
    T *sc;          // u, v, w, T cell values (corrections in DELTA)
    const T *old;   // previous time level of u, v, w, T; null in DELTA
    T *fc;          // faces of u, v, w, p, T
    T *src;
    float *rc;      // [4][N] residuals of the cell equations (RES writes, DELTA reads)
    float *rf;      // [5][NF] residuals of the face equations
    float *TF32;    // RES: FP32 copy of the energy coefficients it builds
    float *grad;    // [9][N] velocity gradients at x0 (RES writes, DELTA reads)
};

// the FP64 state of a context as a View (no residual buffers)
__device__ __host__ inline View<double> view64(const Dev &d)
{
    View<double> v;
    v.gc = d.gc; v.An = d.An; v.AQ = d.AQ; v.wall = d.wall; v.MF = d.MF; v.TF = d.TF;
    v.sc = d.sc; v.old = d.sc + (size_t)S_OLD * d.N; v.fc = d.fc; v.src = d.src;
    v.rc = nullptr; v.rf = nullptr; v.TF32 = nullptr; v.grad = nullptr;
    return v;
}

// One launch of a relaxation sweep: up to four (colour, slab) items, each
// the cells of one colour with k in [k0, k1), given consecutive logical
// blocks starting at bstart. A full-domain colour phase is one item; the
// skewed slab schedule runs colour c on slab s in launch s + 2c.
constexpr int MAX_ITEMS = 4;
struct Sched {
    int n;
    int color[MAX_ITEMS], k0[MAX_ITEMS], k1[MAX_ITEMS];
    int bstart[MAX_ITEMS + 1];
    int rsb;  // one item over all planes, one radial slab per XCD: blocks per slab and plane (0: storage order)
    int rev;  // with rsb: walk the planes from the top
};

// An ensemble launch: per-case device descriptors and, for this launch
// shape, the first block of each case (bstart[ncase] = total blocks)
struct Ens {
    const Dev *dev;
    const View<double> *v64;
    const View<float> *v32;
    const int *bstart;
    int ncase;
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2. The per-iteration kernels renumber them so that XCD x works through
// one contiguous range of cells (a bijection whatever the dispatch order):
// neighbour half-rows, rows and planes are then read through the same L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block()
{
    int b = blockIdx.x, nb = gridDim.x;
    int q = nb / NXCD, rem = nb % NXCD, x = b % NXCD, r = b / NXCD;
    return x * q + (x < rem ? x : rem) + r;
}

// Cell of this thread in the per-cell (source and coefficient) kernels:
//  * storage order: logical block lb covers cells [lb*bs, (lb+1)*bs), so
//    XCD x works through NZ/8 contiguous planes and the planes on either
//    side of its range are read by a second XCD as well;
//  * rslab: XCD x (logical blocks [x*q, (x+1)*q)) works through radial slab
//    x (rows [x*NR/8, (x+1)*NR/8)) plane after plane, so the planes above and
//    below a cell come through its own L2 and only the rows at the slab edges
//    are read twice. Set only when NR % 8 == 0, NT*NR/8 is a multiple of
//    every block size and the grid is N/bs blocks. With krev the planes are
//    walked from the top, so that a kernel starts on the planes the kernel
//    before it ended on, whose data its XCD's L2 still holds;
//  * kfast: blocks numbered plane-fastest: logical block lb covers cells
//    [t*bs, (t+1)*bs) of plane lb % NZ, t = lb / NZ. Returns N for threads
//    past the end of a plane.
__device__ __forceinline__ int cell_of_block(const Dev &d, int lb, int bs)
{
    if (d.rslab) {
        const int sp = d.NT * (d.NR / NXCD), bps = sp / bs, q = d.NZ * bps;
        const int x = lb / q, r = lb - x * q, kk = r / bps, k = d.krev ? d.NZ - 1 - kk : kk;
        return k * (d.NT * d.NR) + x * sp + (r - kk * bps) * bs + (int)threadIdx.x;
    }
    if (!d.kfast) return lb * bs + (int)threadIdx.x;
    const int ppl = d.NT * d.NR;
    const int cp = (lb / d.NZ) * bs + (int)threadIdx.x;
    return cp < ppl ? cp + (lb % d.NZ) * ppl : d.N;
}

// First cell of this thread's pair in the two-cells-per-thread source
// kernels: cell_of_block's map over blocks of 2 bs cells, the thread taking
// the block's storage positions 2t and 2t + 1. The host uses it only
// without kfast and, with rslab, where NT*NR/8 is a multiple of 2 bs.
__device__ __forceinline__ int pair_of_block(const Dev &d, int lb, int bs)
{
    const int bc = 2 * bs;
    if (d.rslab) {
        const int sp = d.NT * (d.NR / NXCD), bps = sp / bc, q = d.NZ * bps;
        const int x = lb / q, r = lb - x * q, kk = r / bps, k = d.krev ? d.NZ - 1 - kk : kk;
        return k * (d.NT * d.NR) + x * sp + (r - kk * bps) * bc + 2 * (int)threadIdx.x;
    }
    return lb * bc + 2 * (int)threadIdx.x;
}

__device__ __forceinline__ int ip1(const Dev &d, int i) { return i + 1 == d.NT ? 0 : i + 1; }
__device__ __forceinline__ int im1(const Dev &d, int i) { return i == 0 ? d.NT - 1 : i - 1; }
// storage position of azimuthal index i within its row: even i first, then
// odd i, so that the cells (and faces) a colour phase touches are contiguous
__device__ __forceinline__ int ipos(const Dev &d, int i) { return (i & 1) * (d.NT >> 1) + (i >> 1); }
__device__ __forceinline__ int iof(const Dev &d, int p)
{
    int h = d.NT >> 1;
    return p < h ? 2 * p : 2 * (p - h) + 1;
}
__device__ __forceinline__ int cidx(const Dev &d, int i, int j, int k) { return ipos(d, i) + d.NT * (j + d.NR * k); }

__device__ __forceinline__ int fidx(const Dev &d, int pos, int i, int j, int k)
{
    switch (pos) {
    case 1: return ipos(d, i) + d.NT * (j + (d.NR + 1) * k);
    case 2: return ipos(d, i) + d.NT * (j + 1 + (d.NR + 1) * k);
    case 3: return d.oT + ipos(d, i) + d.NT * (j + d.NR * k);
    case 4: return d.oT + ipos(d, ip1(d, i)) + d.NT * (j + d.NR * k);
    case 5: return d.oZ + ipos(d, i) + d.NT * (j + d.NR * k);
    default: return d.oZ + ipos(d, i) + d.NT * (j + d.NR * (k + 1));
    }
}

}  // namespace nim
