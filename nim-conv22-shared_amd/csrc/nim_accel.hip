// Anderson acceleration of the outer fixed-point iteration (FP64, gfx950).
//
// No counterpart in the reference: integrate_nse (nse.f90:117-165) repeats
// plain relaxation passes until find_error's increments fall below the
// tolerances. A pass is a map x <- G(x) of the iterate x = [cell values of
// u, v, w, T | faces of u, v, w, p, T]; the converged state is its fixed
// point. The slowest mode of the pressure relaxation decays by about 1e-5 per
// pass at the 1M-cell case, so converging p to 1e-8 takes 1e5 passes or more.
//
// Anderson acceleration (D. G. Anderson 1965; the form of H. F. Walker and
// P. Ni, SIAM J. Numer. Anal. 49 (2011) 1715-1735) treats g(x) = G^d(x), d
// passes (or d mixed-precision cycles), as the map. With f_k = g(x_k) - x_k
// and the differences dF_i = f_{i+1} - f_i, dG_i = g(x_{i+1}) - g(x_i) of the
// last m steps, it takes
//     gamma = argmin || f_k - dF gamma ||,   x_{k+1} = g(x_k) - dG gamma.
// On a linear map this is GMRES on (I - G^d) x = b; the few slow modes that
// stall the plain iteration are removed in a few steps.
//
// The fixed point and the stopping test are those of the plain iteration:
// the test is taken on the increments of each pass (FP64 loop) or on the
// FP64 residual of each cycle (mixed loop), exactly as without acceleration,
// and the solve ends on such a pass or cycle, never on an extrapolated state.
//
// The pressure's constant level is not determined by the equations (Neumann
// walls only), and every pass moves it by the same amount; the least-squares
// problem therefore measures f's pressure faces with their mean removed, so
// that it does not chase that drift.
//
// By default only the pressure faces are extrapolated (nim_accel_scope): the
// slow mode is the pressure's, the other unknowns relax in a few passes once
// p is right, and at a sixth of the iterate's values the extrapolation is
// cheap enough to follow every pass. The other unknowns keep the values of
// the passes.
//
// Device work per acceleration step, one grid-stride pass over the iterate
// each: f and the new difference columns with all dot products of the
// least-squares problem's new Gram row; the update x = g - dG gamma written
// back into the state. The m x m problem is solved on the host; one host
// synchronisation per step reads its new row with the control block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "nim_hip.h"
#include "nim_stencil.h"

using namespace nim;

int32_t nim_internal_fail(int32_t code, const char *msg);
void *nim_internal_stream(nim_ctx *c);
const Dev *nim_internal_dev(nim_ctx *c);
int32_t nim_internal_reset(nim_ctx *c, const double tol[5], int max_iter, int min_iter, bool mixed);
int32_t nim_internal_passes(nim_ctx *c, int n);
int32_t nim_internal_cycles(nim_ctx *c, int n, int inner);
int32_t nim_internal_mixed_prepare(nim_ctx *c);
int32_t nim_internal_read_ctl(nim_ctx *c, const Ctl **out);
void **nim_internal_accel_slot(nim_ctx *c);
void nim_internal_accel_free(void *ws);

namespace {

constexpr int BSX = 256;
constexpr int NBX = 2048;  // blocks of the grid-stride kernels (8 per CU)
constexpr int MAXW = 16;   // largest history (nim_accel_options)

struct Coefs {
    double g[MAXW];
};

// The extrapolated part of the iterate: values [off, off + n) of
// [sc[0, 4N) | fc[0, 5NF)], the pressure faces among them at [p0, p1)
// relative to off. off is 0 (the whole iterate) or lies in fc at an even
// face index (the pressure faces, off = 4N + 3NF, NF even), so that groups
// of four values starting at multiples of 4 stay in one array and aligned.
struct Span {
    long off, n, p0, p1;
};

// address of element s of the iterate [sc[0, 4N) | fc[0, 5NF)] (a Span's
// value s is element off + s)
__device__ __forceinline__ double *it_ptr(const Dev &d, long s)
{
    const long c4 = 4L * d.N;
    return s < c4 ? d.sc + s : d.fc + (s - c4);
}

// the k <= 4 values p[0, k) (one 16-byte load per pair of doubles or four
// floats when k = 4; p is then 16-byte aligned); the rest set to 0
__device__ __forceinline__ void ld4(const double *p, double v[4], int k)
{
    if (k == 4) {
        const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else {
        for (int e = 0; e < 4; e++) v[e] = e < k ? p[e] : 0.;
    }
}

__device__ __forceinline__ void ld4(const float *p, float v[4], int k)
{
    if (k == 4) {
        const float4 a = *(const float4 *)p;
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
        for (int e = 0; e < 4; e++) v[e] = e < k ? p[e] : 0.f;
    }
}

__device__ __forceinline__ void st4(double *p, const double v[4], int k)
{
    if (k == 4) {
        *(double2 *)p = make_double2(v[0], v[1]);
        *(double2 *)(p + 2) = make_double2(v[2], v[3]);
    } else {
        for (int e = 0; e < k; e++) p[e] = v[e];
    }
}

__device__ __forceinline__ void st4(float *p, const float v[4], int k)
{
    if (k == 4)
        *(float4 *)p = make_float4(v[0], v[1], v[2], v[3]);
    else
        for (int e = 0; e < k; e++) p[e] = v[e];
}

// the sum of acc over the block into *out (thread 0 writes)
__device__ __forceinline__ void block_sum(double acc, double *out)
{
    __shared__ double sh[BSX / 64];
    acc = wave_sum(acc);
    __syncthreads();  // sh is reused by consecutive calls
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.;
        for (int q = 0; q < BSX / 64; q++) r += sh[q];
        *out = r;
    }
}

// x = the extrapolated part of the current iterate; one thread per group of
// four values
__global__ void __launch_bounds__(BSX) k_acc_save(Dev d, Span sp, double *x)
{
    const long n = sp.n, ng = (n + 3) >> 2, stride = (long)gridDim.x * BSX;
    for (long gi = (long)blockIdx.x * BSX + threadIdx.x; gi < ng; gi += stride) {
        const long s = gi << 2;
        const int k = (int)(n - s < 4 ? n - s : 4);
        double v[4];
        ld4(it_ptr(d, sp.off + s), v, k);
        st4(x + s, v, k);
    }
}

// out[v] = sum of part[v nb .. v nb + nb), one block per v, fixed order
__global__ void __launch_bounds__(BSX) k_acc_final(const double *part, int nb, double *out)
{
    const double *p = part + (size_t)blockIdx.x * nb;
    double acc = 0.;
    for (int q = threadIdx.x; q < nb; q += BSX) acc += p[q];
    block_sum(acc, out + blockIdx.x);
}

// rows of the partial sums of k_acc_diff
enum { ROW_A = 0, ROW_B = MAXW, ROW_SDF = 2 * MAXW, ROW_SF = 2 * MAXW + 1, NROW = 2 * MAXW + 2 };

// f = g - x, g the current iterate, x the iterate the last d passes started
// from. The previous step's g is x + cprev (cprev = g_prev - x, kept by the
// update), so dG = g - g_prev = f - cprev.
//  * first step (cnt = 0): x = g, fprev = f, cprev = 0;
//  * later steps: the difference columns dF[cur] = f - fprev and
//    dG[cur] = f - cprev, the partial sums of dF[q] . dF[cur] (rows
//    ROW_A + q) and dF[q] . f (ROW_B + q) over the cnt columns in use and of
//    dF[cur] and f over the pressure faces (ROW_SDF, ROW_SF), from which the
//    host takes the products with the pressure faces' mean removed; then
//    fprev = f.
// One thread per group of four values (vector loads of every array).
template <typename VT>
__global__ void __launch_bounds__(BSX) k_acc_diff(Dev d, Span sp, double *x, VT *fprev, VT *cprev, VT *dF, VT *dG,
                                                  long ld, int cur, int cnt, double *part)
{
    const long n = sp.n, ng = (n + 3) >> 2, p0 = sp.p0, p1 = sp.p1;
    const long stride = (long)gridDim.x * BSX;
    double a[MAXW], b[MAXW], sdf = 0., sf = 0.;
#pragma unroll
    for (int q = 0; q < MAXW; q++) a[q] = b[q] = 0.;
    for (long gi = (long)blockIdx.x * BSX + threadIdx.x; gi < ng; gi += stride) {
        const long s = gi << 2;
        const int k = (int)(n - s < 4 ? n - s : 4);
        double g[4], xv[4], f[4];
        ld4(it_ptr(d, sp.off + s), g, k);
        ld4(x + s, xv, k);
        for (int e = 0; e < 4; e++) f[e] = g[e] - xv[e];
        VT fv[4];
        for (int e = 0; e < 4; e++) fv[e] = (VT)f[e];
        if (cnt == 0) {
            const VT z[4] = {0, 0, 0, 0};
            st4(x + s, g, k);
            st4(fprev + s, fv, k);
            st4(cprev + s, z, k);
            continue;
        }
        VT fp[4], cp[4], dfv[4], dgv[4];
        ld4(fprev + s, fp, k);
        ld4(cprev + s, cp, k);
        double df[4];
        for (int e = 0; e < 4; e++) {
            dfv[e] = (VT)(f[e] - (double)fp[e]);
            dgv[e] = (VT)(f[e] - (double)cp[e]);
            df[e] = (double)dfv[e];
            if (s + e >= p0 && s + e < p1) {
                sdf += df[e];
                sf += f[e];
            }
        }
        st4(dF + (size_t)cur * ld + s, dfv, k);
        st4(dG + (size_t)cur * ld + s, dgv, k);
        st4(fprev + s, fv, k);
#pragma unroll
        for (int q = 0; q < MAXW; q++)
            if (q < cnt) {
                VT col[4];
                if (q == cur) {
                    for (int e = 0; e < 4; e++) col[e] = dfv[e];
                } else {
                    ld4(dF + (size_t)q * ld + s, col, k);
                }
                for (int e = 0; e < 4; e++) {
                    a[q] += (double)col[e] * df[e];
                    b[q] += (double)col[e] * f[e];
                }
            }
    }
    if (cnt == 0) return;
    // the block's 2 cnt + 2 sums through one LDS exchange: each wave's sums
    // by shuffles, then thread r adds row r over the waves in wave order
    // (block_sum's order, one barrier instead of two per row)
    __shared__ double sh[BSX / 64][NROW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double v;
#pragma unroll
    for (int q = 0; q < MAXW; q++)
        if (q < cnt) {
            v = wave_sum(a[q]);
            if (lane == 0) sh[wv][ROW_A + q] = v;
            v = wave_sum(b[q]);
            if (lane == 0) sh[wv][ROW_B + q] = v;
        }
    v = wave_sum(sdf);
    if (lane == 0) sh[wv][ROW_SDF] = v;
    v = wave_sum(sf);
    if (lane == 0) sh[wv][ROW_SF] = v;
    __syncthreads();
    const int r = threadIdx.x;
    if ((r < ROW_A + cnt) || (r >= ROW_B && r < ROW_B + cnt) || r == ROW_SDF || r == ROW_SF) {
        double t = 0.;
        for (int w = 0; w < BSX / 64; w++) t += sh[w][r];
        part[(size_t)r * gridDim.x + blockIdx.x] = t;
    }
}

// corr = dG gamma over the cnt columns in use; x = g - corr, written to x
// and the state; cprev = corr (= g - x, for the next step's dG)
template <typename VT>
__global__ void __launch_bounds__(BSX) k_acc_update(Dev d, Span sp, double *x, VT *cprev, const VT *dG, long ld,
                                                    int cnt, Coefs c)
{
    const long n = sp.n, ng = (n + 3) >> 2, stride = (long)gridDim.x * BSX;
    for (long gi = (long)blockIdx.x * BSX + threadIdx.x; gi < ng; gi += stride) {
        const long s = gi << 2;
        const int k = (int)(n - s < 4 ? n - s : 4);
        double *st = it_ptr(d, sp.off + s);
        double g[4], corr[4] = {0., 0., 0., 0.};
        ld4(st, g, k);
#pragma unroll
        for (int q = 0; q < MAXW; q++)
            if (q < cnt) {
                VT col[4];
                ld4(dG + (size_t)q * ld + s, col, k);
                for (int e = 0; e < 4; e++) corr[e] += c.g[q] * (double)col[e];
            }
        double v[4];
        VT cv[4];
        for (int e = 0; e < 4; e++) {
            v[e] = g[e] - corr[e];
            cv[e] = (VT)corr[e];
        }
        st4(x + s, v, k);
        st4(st, v, k);
        st4(cprev + s, cv, k);
    }
}

}  // namespace

// ---- host side --------------------------------------------------------------
struct nim_accel_ws {
    int depth = 12, period = 1, vbytes = 4;  // asked for (nim_accel_options)
    int scope = NIM_ACCEL_PRESSURE;          // extrapolated part (nim_accel_scope)
    int am = 0, avb = 0;                     // allocated for
    long nl = 0;
    double *x = nullptr, *part = nullptr, *out = nullptr;
    void *fprev = nullptr, *cprev = nullptr, *dF = nullptr, *dG = nullptr;  // values of vbytes bytes
    void *block = nullptr;
    double *h_out = nullptr;  // pinned, NROW
    double gram[MAXW][MAXW];  // dF^T dF, pressure faces' mean removed
    double sp[MAXW];          // sum of each column over the pressure faces
};

void nim_internal_accel_free(void *p)
{
    nim_accel_ws *w = (nim_accel_ws *)p;
    if (!w) return;
    if (w->block) (void)hipFree(w->block);
    if (w->h_out) (void)hipHostFree(w->h_out);
    delete w;
}

static nim_accel_ws *accel_ws(nim_ctx *c)
{
    void **slot = nim_internal_accel_slot(c);
    if (!*slot) *slot = new nim_accel_ws();
    return (nim_accel_ws *)*slot;
}

// the extrapolated part of the iterate for a scope
static Span accel_span(const Dev &d, int scope)
{
    const long pf = 4L * d.N + 3L * d.NF;
    if (scope == NIM_ACCEL_PRESSURE) return Span{pf, (long)d.NF, 0, (long)d.NF};
    return Span{0, 4L * d.N + 5L * d.NF, pf, pf + d.NF};
}

// device workspace for the current options: x (FP64), fprev, cprev and
// the 2 m difference columns (vbytes-byte values) over the extrapolated
// part, the partial sums and the results
static int32_t accel_alloc(nim_ctx *c, nim_accel_ws *w)
{
    const Dev *d = nim_internal_dev(c);
    const long nl = accel_span(*d, w->scope).n;
    if (w->block && w->am == w->depth && w->avb == w->vbytes && w->nl == nl) return NIM_OK;
    if (w->block) {
        (void)hipStreamSynchronize((hipStream_t)nim_internal_stream(c));
        (void)hipFree(w->block);
        w->block = nullptr;
    }
    if (!w->h_out && hipHostMalloc((void **)&w->h_out, NROW * sizeof(double)) != hipSuccess) {
        w->h_out = nullptr;
        return nim_internal_fail(NIM_E_HIP, "acceleration host buffer allocation failed");
    }
    // columns padded to 256 bytes
    const long ld = (nl + 63) / 64 * 64;
    const size_t cols = (size_t)(2 * w->depth + 2) * ld * w->vbytes;
    const size_t vecs = (size_t)ld * sizeof(double);
    const size_t red = (size_t)(NROW * NBX + NROW + 64) * sizeof(double);
    void *b = nullptr;
    if (hipMalloc(&b, vecs + cols + red) != hipSuccess)
        return nim_internal_fail(NIM_E_HIP, "acceleration workspace allocation failed");
    w->block = b;
    char *p = (char *)b;
    w->x = (double *)p;
    p += vecs;
    w->fprev = p;
    w->cprev = p + (size_t)ld * w->vbytes;
    w->dF = p + (size_t)2 * ld * w->vbytes;
    w->dG = p + (size_t)(2 + w->depth) * ld * w->vbytes;
    p += cols;
    w->part = (double *)p;
    w->out = w->part + (size_t)NROW * NBX;
    w->am = w->depth;
    w->avb = w->vbytes;
    w->nl = nl;
    return NIM_OK;
}

// Least squares min || f - dF gamma || from the Gram matrix G = dF^T dF and
// r = dF^T f: G scaled to unit diagonal, factored by Cholesky with diagonal
// pivoting (the column with the largest remaining norm first); a column
// whose part outside the span of those before it has a squared norm below
// 1e-14 of its own is dropped (the columns of a converging history become
// nearly dependent), and gamma solves the problem over the columns kept: a
// least-squares minimiser. n^3 / 3 operations, about 1 us at n = 12, while
// the device waits for it after every pass.
static void ls_solve(const double G[MAXW][MAXW], const double *r, int n, double *gamma)
{
    double A[MAXW][MAXW], D[MAXW], y[MAXW];
    int perm[MAXW];
    for (int i = 0; i < n; i++) D[i] = G[i][i] > 0 ? 1. / std::sqrt(G[i][i]) : 0.;
    for (int i = 0; i < n; i++) {
        perm[i] = i;
        for (int j = 0; j < n; j++) A[i][j] = D[i] * G[i][j] * D[j];
    }
    // A = P L L^T P^T over the first `rank` pivots; L in the lower triangle
    int rank = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (A[i][i] > A[p][p]) p = i;
        // columns of unit norm: the remaining diagonal is the squared norm
        // of a column's part outside the span of the pivots before it
        if (!(A[p][p] > 1e-14)) break;
        if (p != k) {
            std::swap(perm[k], perm[p]);
            for (int j = 0; j < n; j++) std::swap(A[k][j], A[p][j]);
            for (int i = 0; i < n; i++) std::swap(A[i][k], A[i][p]);
        }
        const double lkk = std::sqrt(A[k][k]);
        A[k][k] = lkk;
        for (int i = k + 1; i < n; i++) A[i][k] /= lkk;
        for (int i = k + 1; i < n; i++)
            for (int j = k + 1; j <= i; j++) A[i][j] -= A[i][k] * A[j][k];
        for (int i = k + 1; i < n; i++)
            for (int j = i + 1; j < n; j++) A[i][j] = A[j][i];
        rank = k + 1;
    }
    // L L^T z = (D r) permuted, over the kept pivots
    for (int k = 0; k < rank; k++) {
        double t = D[perm[k]] * r[perm[k]];
        for (int j = 0; j < k; j++) t -= A[k][j] * y[j];
        y[k] = t / A[k][k];
    }
    for (int k = rank - 1; k >= 0; k--) {
        double t = y[k];
        for (int j = k + 1; j < rank; j++) t -= A[j][k] * y[j];
        y[k] = t / A[k][k];
    }
    for (int i = 0; i < n; i++) gamma[i] = 0.;
    for (int k = 0; k < rank; k++) gamma[perm[k]] = D[perm[k]] * y[k];
}

// the least-squares solve alone, for host tests: G row-major n x n
extern "C" int32_t nim_internal_ls_solve(const double *G, const double *r, int32_t n, double *gamma)
{
    if (!G || !r || !gamma || n < 1 || n > MAXW) return NIM_E_ARG;
    double g[MAXW][MAXW];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) g[i][j] = G[i * n + j];
    ls_solve(g, r, n, gamma);
    return NIM_OK;
}

static bool verbose()
{
    const char *e = getenv("NIM_VERBOSE");
    return e && e[0] && e[0] != '0';
}

template <typename VT>
static int32_t accel_solve(nim_ctx *c, nim_accel_ws *w, const double tol[5], int32_t max_iter, int32_t min_iter,
                           int32_t inner, int32_t *iters, double err[5])
{
    const Dev &d = *nim_internal_dev(c);
    hipStream_t s = (hipStream_t)nim_internal_stream(c);
    const long ld = (w->nl + 63) / 64 * 64;
    const double inv_nf = 1. / d.NF;
    const Span sp = accel_span(d, w->scope);
    VT *dF = (VT *)w->dF, *dG = (VT *)w->dG, *fprev = (VT *)w->fprev, *cprev = (VT *)w->cprev;
    const int m = w->depth;
    int32_t rc = nim_internal_reset(c, tol, max_iter, min_iter, inner > 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_acc_save, dim3(NBX), dim3(BSX), 0, s, d, sp, w->x);
    // blocks of the difference kernel: each thread takes about GPT groups of
    // four values, so that the block reductions of its 2 m + 2 sums stay
    // small beside the streaming (at most NBX blocks; at the 1M-cell case's
    // pressure faces 389 blocks instead of 2048, 1.7% less time per pass)
    constexpr long GPT = 8;
    const int nbd = (int)std::min<long>(NBX, std::max<long>(64, ((sp.n + 3) / 4 + GPT * BSX - 1) / (GPT * BSX)));
    int cnt = 0, cur = -1, steps = 0;
    bool first = true;
    const Ctl *ctl = nullptr;
    for (;;) {
        rc = inner > 0 ? nim_internal_cycles(c, w->period, inner) : nim_internal_passes(c, w->period);
        if (rc) return rc;
        // f = g - x, the new difference columns and the new row of the
        // least-squares problem, read back with the control block (the state
        // is not touched, so a converged pass or cycle stays as it is)
        const int ncur = first ? -1 : (cur + 1) % m, ncnt = first ? 0 : std::min(cnt + 1, m);
        hipLaunchKernelGGL(k_acc_diff<VT>, dim3(nbd), dim3(BSX), 0, s, d, sp, w->x, fprev, cprev, dF, dG, ld, ncur,
                           ncnt, w->part);
        if (ncnt) {
            hipLaunchKernelGGL(k_acc_final, dim3(NROW), dim3(BSX), 0, s, w->part, nbd, w->out);
            if (hipMemcpyAsync(w->h_out, w->out, NROW * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
                return nim_internal_fail(NIM_E_HIP, "acceleration copy failed");
        }
        if ((rc = nim_internal_read_ctl(c, &ctl))) return rc;
        if (ctl->done) break;
        if (first) {
            first = false;
            continue;
        }
        cur = ncur;
        cnt = ncnt;
        // products with the pressure faces' mean removed:
        // (a - mean a) . (b - mean b) = a . b - sum_p a sum_p b / NF
        const double *h = w->h_out;
        w->sp[cur] = h[ROW_SDF];
        double rhs[MAXW], gam[MAXW];
        bool finite = std::isfinite(h[ROW_SF]);
        for (int q = 0; q < cnt; q++) {
            w->gram[cur][q] = w->gram[q][cur] = h[ROW_A + q] - w->sp[cur] * w->sp[q] * inv_nf;
            rhs[q] = h[ROW_B + q] - w->sp[q] * h[ROW_SF] * inv_nf;
            finite = finite && std::isfinite(w->gram[cur][q]) && std::isfinite(rhs[q]);
        }
        if (!finite) return nim_internal_fail(NIM_E_NONFINITE, "non-finite acceleration history");
        ls_solve(w->gram, rhs, cnt, gam);
        Coefs cf{};
        for (int q = 0; q < cnt; q++) cf.g[q] = gam[q];
        hipLaunchKernelGGL(k_acc_update<VT>, dim3(NBX), dim3(BSX), 0, s, d, sp, w->x, cprev, dG, ld, cnt, cf);
        if (hipGetLastError() != hipSuccess) return nim_internal_fail(NIM_E_HIP, "acceleration launch failed");
        if (verbose() && ++steps % 50 == 0)
            fprintf(stderr, "accel: %d %s, errors %.2e %.2e %.2e %.2e %.2e\n", ctl->iter, inner ? "cycles" : "passes",
                    ctl->err[0], ctl->err[1], ctl->err[2], ctl->err[3], ctl->err[4]);
    }
    if (iters) *iters = ctl->iter;
    if (err)
        for (int q = 0; q < 5; q++) err[q] = ctl->err[q];
    if (ctl->status == 3) return nim_internal_fail(NIM_E_MAXITER, "iteration cap reached");
    if (ctl->status == 4) return nim_internal_fail(NIM_E_NONFINITE, "non-finite increment");
    return NIM_OK;
}

extern "C" int32_t nim_accel_options(nim_ctx *c, int32_t depth, int32_t period, int32_t basis_bytes)
{
    if (!c) return nim_internal_fail(NIM_E_ARG, "null context");
    if (depth < 1 || depth > MAXW || period < 1 || period > 256 || (basis_bytes != 4 && basis_bytes != 8))
        return nim_internal_fail(NIM_E_ARG, "acceleration options: depth in [1, 16], period in [1, 256], "
                                            "history of 4- or 8-byte values");
    nim_accel_ws *w = accel_ws(c);
    w->depth = depth;
    w->period = period;
    w->vbytes = basis_bytes;
    return NIM_OK;
}

extern "C" int32_t nim_accel_scope(nim_ctx *c, int32_t scope)
{
    if (!c) return nim_internal_fail(NIM_E_ARG, "null context");
    if (scope != NIM_ACCEL_ITERATE && scope != NIM_ACCEL_PRESSURE)
        return nim_internal_fail(NIM_E_ARG, "acceleration scope: NIM_ACCEL_ITERATE or NIM_ACCEL_PRESSURE");
    accel_ws(c)->scope = scope;
    return NIM_OK;
}

extern "C" int32_t nim_inner_solve_accel(nim_ctx *c, const double tol[5], int32_t max_iter, int32_t min_iter,
                                         int32_t inner, int32_t *iters, double err[5])
{
    if (!c || !tol) return nim_internal_fail(NIM_E_ARG, "null argument");
    if (max_iter < 1) return nim_internal_fail(NIM_E_ARG, "max_iter must be >= 1");
    if (inner < 0 || inner > 1024) return nim_internal_fail(NIM_E_ARG, "inner passes must be in [0, 1024]");
    int32_t rc;
    if (inner > 0 && (rc = nim_internal_mixed_prepare(c))) return rc;
    nim_accel_ws *w = accel_ws(c);
    if ((rc = accel_alloc(c, w))) return rc;
    return w->vbytes == 8 ? accel_solve<double>(c, w, tol, max_iter, min_iter, inner, iters, err)
                          : accel_solve<float>(c, w, tol, max_iter, min_iter, inner, iters, err);
}
