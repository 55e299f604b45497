// Device helpers shared by the time-step kernels and their adjoint: wave
// reductions, the colour-phase cell enumeration, the finite differences of
// the face fields and the NIM exponential fits (templated on the scalar, so
// that the adjoint differentiates the same expressions with dual numbers).
#pragma once
#include "nim_kernels.h"

namespace nim {

// ---------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------
// running maximum of |x| for a convergence measure: a non-finite x (NaN
// included, which fmax would drop) counts as infinity, so it is reported
__device__ __forceinline__ double err_max(double e, double x)
{
    const double a = fabs(x);
    return fmax(e, a <= DBL_MAX ? a : INFINITY);
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// cell of colour `color` with k in [k0, k1) handled by thread `tid`
__device__ __forceinline__ bool color_cell(const Dev &d, int color, int k0, int k1, int tid, int &i, int &j, int &k)
{
    int nth = d.NT >> 1, nrh = (d.NR + 1) >> 1;
    int ih = tid % nth, rest = tid / nth;
    int jj = rest % nrh;
    k = k0 + rest / nrh;
    if (k >= k1) return false;
    j = ((color + k) & 1) + 2 * jj;
    if (j >= d.NR) return false;
    int r = (color - j - 3 * k) & 3;
    i = 2 * ih + (r >> 1);
    return true;
}

// ---------------------------------------------------------------------
// finite differences shared by sources, sweeps and wall updates
// ---------------------------------------------------------------------
// derivative across cells of the face value at position fpos, along
// direction dir (0 theta, 1 r, 2 z): central (f(+1) - f(-1))/4, 2nd-order
// one-sided at walls (nse.f90:1643-1653 and its repetitions)
//
// The stencil points and their weights are chosen by selects, so the loads
// of all derivatives of a cell issue together instead of one branch (and
// one memory round trip) at a time. Weights (1, -1, 0) give the central
// difference exactly (the products by +-1 and 0 are exact); (-3, 4, -1) and
// (3, -4, 1) the one-sided forms in the reference's operation order. The
// third point is loaded only at a wall: a wave's cells share j and k (half
// an azimuthal row), so that condition is uniform over the wave, and the
// central rows (all but 2 of n_r and of n_z) issue two loads, not three.
// Its weight is then 0 and it enters as +0, which leaves every sum's bits
// as the product 0 * f(-1) did (the FP64 loop's fields are bitwise those of
// three loads; in FP32 the compiler's contractions into fused multiply-adds
// shift with the code and move the last bit).
template <typename T>
__device__ __forceinline__ T three_point(T f0, T f1, T f2, bool lo, bool hi)
{
    T w0 = lo ? T(-3) : hi ? T(3) : T(1);
    T w1 = lo ? T(4) : hi ? T(-4) : T(-1);
    T w2 = lo ? T(-1) : hi ? T(1) : T(0);
    return (w0 * f0 + w1 * f1 + w2 * f2) / T(4);
}

template <typename T>
__device__ __forceinline__ T tder(const Dev &d, const T *F, int fpos, int i, int j, int k, int dir)
{
    if (dir == 0) return (F[fidx(d, fpos, ip1(d, i), j, k)] - F[fidx(d, fpos, im1(d, i), j, k)]) / T(4);
    if (dir == 1) {
        bool lo = j == 0, hi = j == d.NR - 1;
        // central: (j+1, j-1); low wall: (0, 1, 2); high wall: (j, j-1, j-2)
        int j0 = lo ? 0 : hi ? j : j + 1, j1 = lo ? 1 : j - 1, j2 = lo ? 2 : j - 2;
        const T f0 = F[fidx(d, fpos, i, j0, k)], f1 = F[fidx(d, fpos, i, j1, k)];
        const T f2 = (lo || hi) ? F[fidx(d, fpos, i, j2, k)] : T(0);
        return three_point(f0, f1, f2, lo, hi);
    }
    bool lo = k == 0, hi = k == d.NZ - 1;
    int k0 = lo ? 0 : hi ? k : k + 1, k1 = lo ? 1 : k - 1, k2 = lo ? 2 : k - 2;
    const T f0 = F[fidx(d, fpos, i, j, k0)], f1 = F[fidx(d, fpos, i, j, k1)];
    const T f2 = (lo || hi) ? F[fidx(d, fpos, i, j, k2)] : T(0);
    return three_point(f0, f1, f2, lo, hi);
}

// the cross-diffusion combination of the nine derivatives of one cell,
// with the cell's G (3) and GO (6) metrics given or loaded
template <typename T>
__device__ __forceinline__ T cross_combine(T coef, const T G[3], const T GO[6], T fx, T fy, T fz, T fxy, T fyx, T fxz,
                                           T fzx, T fyz, T fzy)
{
    return coef * (G[0] * fx + G[1] * fy + G[2] * fz - GO[0] * fxy - GO[1] * fyx - GO[2] * fxz - GO[3] * fzx
                   - GO[4] * fyz - GO[5] * fzy);
}

template <typename T>
__device__ __forceinline__ void cross_metrics(const Dev &d, const T *g, int c, T G[3], T GO[6])
{
    const int N = d.N;
#pragma unroll
    for (int m = 0; m < 3; m++) G[m] = g[(G_G + m) * N + c];
#pragma unroll
    for (int m = 0; m < 6; m++) GO[m] = g[(G_GO + m) * N + c];
}

template <typename T>
__device__ __forceinline__ T cross_combine(const Dev &d, const T *g, int c, T coef, T fx, T fy, T fz, T fxy, T fyx,
                                           T fxz, T fzx, T fyz, T fzy)
{
    T G[3], GO[6];
    cross_metrics(d, g, c, G, GO);
    return cross_combine(coef, G, GO, fx, fy, fz, fxy, fyx, fxz, fzx, fyz, fzy);
}

// cross-diffusion operator of calculate_R_D2 / _R_D_p2 / _R_D_T
// (nse.f90:1628-2134) with the cell's metrics G, GO already loaded
template <typename T>
__device__ __forceinline__ T cross_diffusion(const Dev &d, const T G[3], const T GO[6], const T *F, int i, int j,
                                             int k, T coef)
{
    T fxy = T(0.5) * (tder(d, F, 2, i, j, k, 0) - tder(d, F, 1, i, j, k, 0));
    T fzy = T(0.5) * (tder(d, F, 6, i, j, k, 0) - tder(d, F, 5, i, j, k, 0));
    T fyx = T(0.5) * (tder(d, F, 4, i, j, k, 1) - tder(d, F, 3, i, j, k, 1));
    T fzx = T(0.5) * (tder(d, F, 6, i, j, k, 1) - tder(d, F, 5, i, j, k, 1));
    T fxz = T(0.5) * (tder(d, F, 2, i, j, k, 2) - tder(d, F, 1, i, j, k, 2));
    T fyz = T(0.5) * (tder(d, F, 4, i, j, k, 2) - tder(d, F, 3, i, j, k, 2));
    T fx = (F[fidx(d, 2, i, j, k)] - F[fidx(d, 1, i, j, k)]) / T(2);
    T fy = (F[fidx(d, 4, i, j, k)] - F[fidx(d, 3, i, j, k)]) / T(2);
    T fz = (F[fidx(d, 6, i, j, k)] - F[fidx(d, 5, i, j, k)]) / T(2);
    return cross_combine(coef, G, GO, fx, fy, fz, fxy, fyx, fxz, fzx, fyz, fzy);
}

// the same, loading the metrics; g is the per-cell geometry block [NGC][N]
template <typename T>
__device__ T cross_diffusion(const Dev &d, const T *g, const T *F, int i, int j, int k, int c, T coef)
{
    T fxy = T(0.5) * (tder(d, F, 2, i, j, k, 0) - tder(d, F, 1, i, j, k, 0));
    T fzy = T(0.5) * (tder(d, F, 6, i, j, k, 0) - tder(d, F, 5, i, j, k, 0));
    T fyx = T(0.5) * (tder(d, F, 4, i, j, k, 1) - tder(d, F, 3, i, j, k, 1));
    T fzx = T(0.5) * (tder(d, F, 6, i, j, k, 1) - tder(d, F, 5, i, j, k, 1));
    T fxz = T(0.5) * (tder(d, F, 2, i, j, k, 2) - tder(d, F, 1, i, j, k, 2));
    T fyz = T(0.5) * (tder(d, F, 4, i, j, k, 2) - tder(d, F, 3, i, j, k, 2));
    T fx = (F[fidx(d, 2, i, j, k)] - F[fidx(d, 1, i, j, k)]) / T(2);
    T fy = (F[fidx(d, 4, i, j, k)] - F[fidx(d, 3, i, j, k)]) / T(2);
    T fz = (F[fidx(d, 6, i, j, k)] - F[fidx(d, 5, i, j, k)]) / T(2);
    return cross_combine(d, g, c, coef, fx, fy, fz, fxy, fyx, fxz, fzx, fyz, fzy);
}

// ---------------------------------------------------------------------
// LDS-staged tiles of one plane for the source kernels
// ---------------------------------------------------------------------
// A block of 256 threads takes a tile of TI x TJ cells (TI consecutive
// azimuthal indices, TJ = 256 / TI radial rows) of one plane k and stages in
// LDS, with one coalesced pass, the faces of that plane its cells' stencils
// read: radial faces of rows j0 .. j0+TJ, azimuthal faces of rows j0-2 ..
// j0+TJ+1 (the one-sided wall closures reach two rows) and axial faces of
// planes k and k+1, each with the theta neighbours. The derivatives along
// theta and r and the cell's own face differences then come from LDS; the
// ones along z (other planes) from global memory. The arithmetic is that of
// cross_diffusion, term by term, so the results are the same bits.
template <int TI>
struct FaceTile {
    static constexpr int TJ = 256 / TI;
    double R[TJ + 1][TI + 2];     // R(i0-1+li, j0+jj, k)
    double Tf[TJ + 4][TI + 1];    // T(i0+li, j0-2+jj, k)
    double Z[2][TJ + 4][TI + 2];  // Z(i0-1+li, j0-2+jj, k+kk)
};

__device__ __forceinline__ int wrap_i(const Dev &d, int i) { return i < 0 ? i + d.NT : i >= d.NT ? i - d.NT : i; }

template <int TI>
__device__ __forceinline__ void load_tile(const Dev &d, const double *F, FaceTile<TI> &s, int i0, int j0, int k)
{
    constexpr int TJ = FaceTile<TI>::TJ;
    const int NR = d.NR;
    for (int e = threadIdx.x; e < (TJ + 1) * (TI + 2); e += blockDim.x) {
        const int jj = e / (TI + 2), li = e - jj * (TI + 2), j = j0 + jj;
        if (j <= NR) s.R[jj][li] = F[fidx(d, 1, wrap_i(d, i0 - 1 + li), j, k)];
    }
    for (int e = threadIdx.x; e < (TJ + 4) * (TI + 1); e += blockDim.x) {
        const int jj = e / (TI + 1), li = e - jj * (TI + 1), j = j0 - 2 + jj;
        if (j >= 0 && j < NR) s.Tf[jj][li] = F[fidx(d, 3, wrap_i(d, i0 + li), j, k)];
    }
    for (int e = threadIdx.x; e < 2 * (TJ + 4) * (TI + 2); e += blockDim.x) {
        const int kk = e / ((TJ + 4) * (TI + 2)), r = e - kk * ((TJ + 4) * (TI + 2));
        const int jj = r / (TI + 2), li = r - jj * (TI + 2), j = j0 - 2 + jj;
        if (j >= 0 && j < NR) s.Z[kk][jj][li] = F[fidx(d, 5, wrap_i(d, i0 - 1 + li), j, k + kk)];
    }
}

// cross_diffusion of cell (i, j, k) = (i0 + li, j0 + jr, k) from a staged tile
template <int TI>
__device__ __forceinline__ double cross_diffusion_tile(const Dev &d, const FaceTile<TI> &s, const double *F, int i,
                                                       int j, int k, int c, int li, int jr, int j0, double coef)
{
    // rows of the r-derivative's three points (tder, dir 1), as tile rows of Tf and Z
    const bool lo = j == 0, hi = j == d.NR - 1;
    const int ja = (lo ? 0 : hi ? j : j + 1) - j0 + 2, jb = (lo ? 1 : j - 1) - j0 + 2;
    const int jc = (lo ? 2 : hi ? j - 2 : j - 1) - j0 + 2, jm = jr + 2;
    const double tR2 = (s.R[jr + 1][li + 2] - s.R[jr + 1][li]) / 4.;
    const double tR1 = (s.R[jr][li + 2] - s.R[jr][li]) / 4.;
    const double tZ6 = (s.Z[1][jm][li + 2] - s.Z[1][jm][li]) / 4.;
    const double tZ5 = (s.Z[0][jm][li + 2] - s.Z[0][jm][li]) / 4.;
    const double fxy = 0.5 * (tR2 - tR1);
    const double fzy = 0.5 * (tZ6 - tZ5);
    const double fyx = 0.5 * (three_point(s.Tf[ja][li + 1], s.Tf[jb][li + 1], s.Tf[jc][li + 1], lo, hi) -
                              three_point(s.Tf[ja][li], s.Tf[jb][li], s.Tf[jc][li], lo, hi));
    const double fzx = 0.5 * (three_point(s.Z[1][ja][li + 1], s.Z[1][jb][li + 1], s.Z[1][jc][li + 1], lo, hi) -
                              three_point(s.Z[0][ja][li + 1], s.Z[0][jb][li + 1], s.Z[0][jc][li + 1], lo, hi));
    const double fxz = 0.5 * (tder(d, F, 2, i, j, k, 2) - tder(d, F, 1, i, j, k, 2));
    const double fyz = 0.5 * (tder(d, F, 4, i, j, k, 2) - tder(d, F, 3, i, j, k, 2));
    const double fx = (s.R[jr + 1][li + 1] - s.R[jr][li + 1]) / 2.;
    const double fy = (s.Tf[jm][li + 1] - s.Tf[jm][li]) / 2.;
    const double fz = (s.Z[1][jm][li + 1] - s.Z[0][jm][li + 1]) / 2.;
    return cross_combine(d, d.gc, c, coef, fx, fy, fz, fxy, fyx, fxz, fzx, fyz, fzy);
}

// tile of block lb (a bijection onto the NT/TI x ceil(NR/TJ) x NZ tiles);
// the grid is tile_blocks of them
template <int TI>
__device__ __forceinline__ void tile_of_block(const Dev &d, int lb, int &i0, int &j0, int &k)
{
    constexpr int TJ = FaceTile<TI>::TJ;
    const int nti = d.NT / TI, ntj = (d.NR + TJ - 1) / TJ;
    k = lb / (nti * ntj);
    const int r = lb - k * nti * ntj, tj = r / nti;
    i0 = (r - tj * nti) * TI;
    j0 = tj * TJ;
}

// gradient component of a face field by the divergence theorem over the six
// faces (dphi_dx2..dphi_dz2, nse.f90:2866-2903); An is [3][NF]
template <typename T>
__device__ __forceinline__ T dphi(const Dev &d, const T *An, const T phi[6], const int f[6], int comp, T invV)
{
    const T *A = An + (size_t)comp * d.NF;
    T s = phi[1] * A[f[1]] - phi[0] * A[f[0]] + phi[3] * A[f[3]] - phi[2] * A[f[2]] + phi[5] * A[f[5]]
          - phi[4] * A[f[4]];
    return invV * s;
}

__device__ __forceinline__ void cell_faces(const Dev &d, int i, int j, int k, int f[6])
{
    for (int p = 0; p < 6; p++) f[p] = fidx(d, p + 1, i, j, k);
}

// ---------------------------------------------------------------------
// two cells per thread (source kernels, NIM_PAIR)
// ---------------------------------------------------------------------
// A thread takes storage positions q and q + 1 (q even) of one half-row,
// the azimuthal indices i and i + 2: every field value of the two cells at
// the same offset is one aligned 16-byte (FP64) or 8-byte (FP32) load. Their
// azimuthal neighbours i - 1, i + 1, i + 3 are three consecutive positions
// of the other half-row, covered by two aligned pairs. Needs n_theta % 4 ==
// 0 (even half-rows) and even block offsets, which n_theta % 4 == 0 gives.
template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <typename T>
__device__ __forceinline__ void ld2(const T *p, T &a, T &b)
{
    const typename Vec2<T>::type v = *reinterpret_cast<const typename Vec2<T>::type *>(p);
    a = v.x;
    b = v.y;
}

template <typename T>
__device__ __forceinline__ void st2(T *p, T a, T b)
{
    typename Vec2<T>::type v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<typename Vec2<T>::type *>(p) = v;
}

struct CellPair {
    int i, j, k;  // indices of the first cell (the second is i + 2)
    int q;        // storage position of the first cell in its row
    int lo, hi;   // other-half positions of the pairs holding i - 1, i + 1, i + 3
    bool odd;     // the pair lies in the odd half-row
};

// c is the first cell of the pair (an even storage position)
__device__ __forceinline__ CellPair cell_pair(const Dev &d, int c)
{
    CellPair p;
    const int h = d.NT >> 1;
    p.q = c % d.NT;
    p.j = (c / d.NT) % d.NR;
    p.k = c / (d.NT * d.NR);
    p.i = iof(d, p.q);
    p.odd = p.q >= h;
    const int qq = p.odd ? p.q - h : p.q, ob = p.odd ? 0 : h;
    // even half, i = 2qq: i - 1 sits at ob + qq - 1 (the second of the pair
    // before ob + qq), i + 1 and i + 3 at ob + qq and ob + qq + 1;
    // odd half, i = 2qq + 1: i - 1, i + 1 at ob + qq, ob + qq + 1, i + 3 at
    // ob + qq + 2; both wrap within the half-row
    p.lo = ob + (p.odd ? qq : (qq - 2 + h) % h);
    p.hi = ob + (p.odd ? (qq + 2) % h : qq);
    return p;
}

// first storage index of row (j, k) of face position pos (1..6; 4 shares
// the azimuthal block of 3) or of the cells (pos 0)
__device__ __forceinline__ int row_base(const Dev &d, int pos, int j, int k)
{
    switch (pos) {
    case 0: return d.NT * (j + d.NR * k);
    case 1: return d.NT * (j + (d.NR + 1) * k);
    case 2: return d.NT * (j + 1 + (d.NR + 1) * k);
    case 3:
    case 4: return d.oT + d.NT * (j + d.NR * k);
    case 5: return d.oZ + d.NT * (j + d.NR * k);
    default: return d.oZ + d.NT * (j + d.NR * (k + 1));
    }
}

// values of row rb at the azimuthal neighbours i - 1, i + 1, i + 3
template <typename T>
__device__ __forceinline__ void other3(const CellPair &p, const T *F, int rb, T &m1, T &p1, T &p3)
{
    T la, lb, ha, hb;
    ld2(F + rb + p.lo, la, lb);
    ld2(F + rb + p.hi, ha, hb);
    m1 = p.odd ? la : lb;
    p1 = p.odd ? lb : ha;
    p3 = p.odd ? ha : hb;
}

// face value at position pos (1..6) of row (j, k) for both cells
template <typename T>
__device__ __forceinline__ void face2(const Dev &d, const CellPair &p, const T *F, int pos, int j, int k, T &a, T &b)
{
    if (pos == 4) {
        T m1;
        other3(p, F, row_base(d, 4, j, k), m1, a, b);
        return;
    }
    ld2(F + row_base(d, pos, j, k) + p.q, a, b);
}

// tder for both cells (same arithmetic per cell)
template <typename T>
__device__ __forceinline__ void tder2(const Dev &d, const CellPair &p, const T *F, int fpos, int dir, T &a, T &b)
{
    const int j = p.j, k = p.k;
    if (dir == 0) {
        T m1, p1, p3;
        other3(p, F, row_base(d, fpos, j, k), m1, p1, p3);
        a = (p1 - m1) / T(4);
        b = (p3 - p1) / T(4);
        return;
    }
    T f0a, f0b, f1a, f1b, f2a = T(0), f2b = T(0);
    bool lo, hi;
    if (dir == 1) {
        lo = j == 0; hi = j == d.NR - 1;
        const int j0 = lo ? 0 : hi ? j : j + 1, j1 = lo ? 1 : j - 1, j2 = lo ? 2 : j - 2;
        face2(d, p, F, fpos, j0, k, f0a, f0b);
        face2(d, p, F, fpos, j1, k, f1a, f1b);
        if (lo || hi) face2(d, p, F, fpos, j2, k, f2a, f2b);
    } else {
        lo = k == 0; hi = k == d.NZ - 1;
        const int k0 = lo ? 0 : hi ? k : k + 1, k1 = lo ? 1 : k - 1, k2 = lo ? 2 : k - 2;
        face2(d, p, F, fpos, j, k0, f0a, f0b);
        face2(d, p, F, fpos, j, k1, f1a, f1b);
        if (lo || hi) face2(d, p, F, fpos, j, k2, f2a, f2b);
    }
    a = three_point(f0a, f1a, f2a, lo, hi);
    b = three_point(f0b, f1b, f2b, lo, hi);
}

// cross_diffusion of both cells, each with its metrics (G, GO)
template <typename T>
__device__ __forceinline__ void cross_diffusion2(const Dev &d, const CellPair &p, const T *F, const T GA[3],
                                                 const T GOA[6], const T GB[3], const T GOB[6], T coef, T &ra, T &rb)
{
    T t2a, t2b, t1a, t1b, t6a, t6b, t5a, t5b;
    tder2(d, p, F, 2, 0, t2a, t2b);
    tder2(d, p, F, 1, 0, t1a, t1b);
    tder2(d, p, F, 6, 0, t6a, t6b);
    tder2(d, p, F, 5, 0, t5a, t5b);
    const T fxya = T(0.5) * (t2a - t1a), fxyb = T(0.5) * (t2b - t1b);
    const T fzya = T(0.5) * (t6a - t5a), fzyb = T(0.5) * (t6b - t5b);
    tder2(d, p, F, 4, 1, t2a, t2b);
    tder2(d, p, F, 3, 1, t1a, t1b);
    tder2(d, p, F, 6, 1, t6a, t6b);
    tder2(d, p, F, 5, 1, t5a, t5b);
    const T fyxa = T(0.5) * (t2a - t1a), fyxb = T(0.5) * (t2b - t1b);
    const T fzxa = T(0.5) * (t6a - t5a), fzxb = T(0.5) * (t6b - t5b);
    tder2(d, p, F, 2, 2, t2a, t2b);
    tder2(d, p, F, 1, 2, t1a, t1b);
    tder2(d, p, F, 4, 2, t6a, t6b);
    tder2(d, p, F, 3, 2, t5a, t5b);
    const T fxza = T(0.5) * (t2a - t1a), fxzb = T(0.5) * (t2b - t1b);
    const T fyza = T(0.5) * (t6a - t5a), fyzb = T(0.5) * (t6b - t5b);
    T fa[6], fb[6];
#pragma unroll
    for (int q = 0; q < 6; q++) face2(d, p, F, q + 1, p.j, p.k, fa[q], fb[q]);
    const T fxa = (fa[1] - fa[0]) / T(2), fxb = (fb[1] - fb[0]) / T(2);
    const T fya = (fa[3] - fa[2]) / T(2), fyb = (fb[3] - fb[2]) / T(2);
    const T fza = (fa[5] - fa[4]) / T(2), fzb = (fb[5] - fb[4]) / T(2);
    ra = cross_combine(coef, GA, GOA, fxa, fya, fza, fxya, fyxa, fxza, fzxa, fyza, fzya);
    rb = cross_combine(coef, GB, GOB, fxb, fyb, fzb, fxyb, fyxb, fxzb, fzxb, fyzb, fzyb);
}

// the cross metrics of both cells (c the first)
template <typename T>
__device__ __forceinline__ void cross_metrics2(const Dev &d, const T *g, int c, T GA[3], T GOA[6], T GB[3], T GOB[6])
{
    const int N = d.N;
#pragma unroll
    for (int m = 0; m < 3; m++) ld2(g + (size_t)(G_G + m) * N + c, GA[m], GB[m]);
#pragma unroll
    for (int m = 0; m < 6; m++) ld2(g + (size_t)(G_GO + m) * N + c, GOA[m], GOB[m]);
}

// the six face values of a face field for both cells
template <typename T>
__device__ __forceinline__ void faces2(const Dev &d, const CellPair &p, const T *F, T a[6], T b[6])
{
#pragma unroll
    for (int q = 0; q < 6; q++) face2(d, p, F, q + 1, p.j, p.k, a[q], b[q]);
}

// dphi from loaded values: the face values phi and area components A of one cell
template <typename T>
__device__ __forceinline__ T dphi_vals(const T phi[6], const T A[6], T invV)
{
    T s = phi[1] * A[1] - phi[0] * A[0] + phi[3] * A[3] - phi[2] * A[2] + phi[5] * A[5] - phi[4] * A[4];
    return invV * s;
}

// ---------------------------------------------------------------------
// forward-mode dual numbers (value, derivative)
// ---------------------------------------------------------------------
struct Dual {
    double v, d;
};
__device__ __forceinline__ double dval(double x) { return x; }
__device__ __forceinline__ double dval(Dual x) { return x.v; }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ Dual dexp(Dual x)
{
    double e = exp(x.v);
    return {e, e * x.d};
}
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b)
{
    double q = a / b.v;
    return {q, -q * b.d / b.v};
}

// ---------------------------------------------------------------------
// NIM exponential-fit constants of one direction (nim.f90:184-287)
// ---------------------------------------------------------------------
// Both branches are the reference's expressions rewritten with the
// reciprocals rk = 1/k, ru = 1/u and t = u/k (= Re/2): the Taylor branch
// becomes series in t with exact rational coefficients (Horner form, one
// division instead of about forty), the exponential branch uses
// p = 1/(e^-Re - 1) and q = 1/(1 - e^Re) (five divisions instead of
// eleven). The values agree with the reference's to rounding, amplified
// only by the cancellation the closed forms already have near |Re| =
// v_limit.
// R is double, or Dual for the derivatives the adjoint needs.
template <typename R>
__device__ __forceinline__ void nim_direction_t(R u, R k, double gp, double gm, double vlim, R A[6])
{
    const R rk = 1. / k;
    const R t = u * rk, Re = 2. * t;
    if (fabs(dval(Re)) > vlim) {
        const R ru = 1. / u;
        const R e = dexp(Re), ie = 1. / e;
        const R p = 1. / (ie - 1.), q = 1. / (1. - e);
        A[0] = (2. * rk * p + ru) * gp;
        A[1] = (-(2. * rk * q) - ru) * gm;
        A[2] = (u * rk * p) * gp;
        A[3] = (u * rk * q) * gm;
        if (dval(Re) > 0) {
            A[4] = (-k - u * (ie + 1.) * p) * (ru * ru);
            A[5] = -p - 0.5 * k * ru;
        } else {
            A[4] = (-k - u * (1. + e) * q) * (ru * ru);
            A[5] = -(e * q) - 0.5 * k * ru;
        }
    } else {
        constexpr double c3 = 1. / 3., c6 = 1. / 6., c45 = 1. / 45., c90 = 1. / 90., c945 = 1. / 945.;
        constexpr double c945b = 2. / 945., c4725 = 1. / 4725., c9450 = 1. / 9450., c93555 = 1. / 93555.;
        constexpr double c93555b = 2. / 93555., c11 = 1382. / 6.38512875e8;
        const R s2 = t * t;
        // odd and even series shared by the pairs A0/A1 and A2/A3
        const R Q = t * (c3 + s2 * (-c45 + s2 * (c945b + s2 * (c4725 + s2 * c93555b))));
        const R E = s2 * (-c6 + s2 * (c90 + s2 * (c945 + s2 * (c9450 - s2 * c93555))));
        A[0] = (rk * (-1. - Q)) * gp;
        A[1] = (rk * (-1. + Q)) * gm;
        A[2] = (-0.5 - 0.5 * t + E) * gp;
        A[3] = (-0.5 + 0.5 * t + E) * gm;
        A[4] = rk * (c3 + s2 * (-c45 + s2 * (c945b + s2 * (c4725 + s2 * (c93555b + s2 * c11)))));
        A[5] = 0.5 + t * (c6 + s2 * (-c90 + s2 * (c945 + s2 * (c9450 + s2 * c93555))));
    }
}


__device__ void nim_direction(double u, double k, double gp, double gm, double vlim, double A[6])
{
    nim_direction_t<double>(u, k, gp, gm, vlim, A);
}

}  // namespace nim
