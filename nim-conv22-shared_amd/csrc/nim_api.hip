// Host side of libnim_hip.so: the C-ABI of include/nim_hip.h, layout
// conversion between the reference's numbering and the device layout, and
// the launch sequence of one outer iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nim_hip.h"
#include "nim_kernels.h"

namespace nim {
__global__ void k_coef_A(Dev d, int energy);
__global__ void k_coef_F(Dev d, int energy);
__global__ void k_T_coef(Dev d);
__global__ void k_mom_src(Dev d);
__global__ void k_ppe_src(Dev d);
__global__ void k_T_src(Dev d);
__global__ void k_mom_src2(Dev d);
__global__ void k_ppe_src2(Dev d);
__global__ void k_T_src2(Dev d);
__global__ void k_res_ppe_src2(Dev d, View<double> v);
template <int TI> __global__ void k_T_src_tile(Dev d);
template <int TI> __global__ void k_mom_src_tile(Dev d);
template <int TI> __global__ void k_ppe_src_tile(Dev d);
template <bool NTL> __global__ void k_mom_sweep(Dev d, Sched sc);
__global__ void k_T_sweep(Dev d, Sched sc);
template <bool NTL> __global__ void k_T_sweep_fused(Dev d, Sched sc);
__global__ void k_zero_flux(Dev d);
__global__ void k_energy_coefs(Dev d);
template <bool NTL> __global__ void k_ppe_sweep(Dev d, Sched sc);
__global__ void k_solve_reset(Dev d, double t0, double t1, double t2, double t3, double t4, int max_iter, int min_iter,
                              int p_measure);
__global__ void k_check(Dev d);
__global__ void k_pressure_coeffs(Dev d);
__global__ void k_wall_pressure_coeffs(Dev d);
__global__ void k_advance(Dev d);
__global__ void k_end_step_partial(Dev d);
__global__ void k_end_step_final(Dev d, int nblocks, double *out);
__global__ void k_nusselt(Dev d, double R_i, double R_o, double *out);
__global__ void k_res_mom(Dev d, View<double> v);
__global__ void k_res_ppe_src(Dev d, View<double> v);
__global__ void k_res_ppe(Dev d, View<double> v);
__global__ void k_res_T_coef(Dev d, View<double> v);
__global__ void k_res_T(Dev d, View<double> v);
__global__ void k_d_mom_src(Dev d, View<float> v);
template <bool NTL> __global__ void k_d_mom_sweep(Dev d, View<float> v, Sched sc);
__global__ void k_d_ppe_src(Dev d, View<float> v);
template <bool NTL> __global__ void k_d_ppe_sweep(Dev d, View<float> v, Sched sc);
__global__ void k_d_zero_flux(Dev d, View<float> v);
__global__ void k_d_T_src(Dev d, View<float> v);
template <bool NTL> __global__ void k_d_T_sweep(Dev d, View<float> v, Sched sc);
__global__ void k_mix_update(Dev d, View<double> x, View<float> dx, double inv_inner);
__global__ void k_cast(const double *in, float *out, long n);
__global__ void k_e_mom_src(Ens e);
__global__ void k_e_mom_sweep(Ens e, int color);
__global__ void k_e_ppe_src(Ens e);
__global__ void k_e_ppe_sweep(Ens e, int color);
__global__ void k_e_T_coef(Ens e);
__global__ void k_e_T_src(Ens e);
__global__ void k_e_T_sweep(Ens e, int color);
__global__ void k_e_check(Ens e);
__global__ void k_e_res_mom(Ens e);
__global__ void k_e_res_ppe_src(Ens e);
__global__ void k_e_res_ppe(Ens e);
__global__ void k_e_res_T_coef(Ens e);
__global__ void k_e_res_T(Ens e);
__global__ void k_e_d_mom_src(Ens e);
__global__ void k_e_d_mom_sweep(Ens e, int color);
__global__ void k_e_d_ppe_src(Ens e);
__global__ void k_e_d_ppe_sweep(Ens e, int color);
__global__ void k_e_d_zero_flux(Ens e);
__global__ void k_e_d_T_src(Ens e);
__global__ void k_e_d_T_sweep(Ens e, int color);
__global__ void k_e_mix_update(Ens e, double inv_inner);
__global__ void k_e_gather(Ens e, Ctl *out);
__global__ void k_e_coef_A(Ens e);
__global__ void k_e_coef_F(Ens e, int cast);
__global__ void k_e_end_step_partial(Ens e);
__global__ void k_e_end_step_final(Ens e, const int *nred, double *out);
__global__ void k_e_advance(Ens e);
__global__ void k_e_nusselt(Ens e, const double *R, double *out);
}  // namespace nim

using namespace nim;

static thread_local std::string g_err;

static int32_t fail(int32_t code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess)                                                                           \
            return fail(NIM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));                      \
    } while (0)

static constexpr int BS = 256;
static constexpr int NT_MIN_CELLS = 1 << 19;  // see nim_ctx::nt

// Scratch of nim_end_step_noalloc: the per-block partial results of the
// end-of-step reductions (6 doubles per block of 256 cells), in a
// file-scope device array instead of the context's allocation, so that the
// entry point allocates nothing. Each device holds its own copy for the
// life of the process; every call on a device shares it.
static constexpr int NOALLOC_MAX_CELLS = 1 << 24;
static constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)6 * (NOALLOC_MAX_CELLS / 256);
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

// Device address of `scratch` on the current device, resolved once per
// device (hipGetSymbolAddress with that device current) and cached. The
// array's name is only a host-side stand-in; host code uses this address.
static double *noalloc_scratch_current_device()
{
    static std::mutex mu;
    static std::vector<double *> ptrs;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= ptrs.size()) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs[dev] = (double *)p;
    }
    return ptrs[dev];
}

struct nim_ctx {
    nim_dims dims;
    nim_params prm;
    int NT, NR, NZ, N, NF, NB, oT, oZ;
    hipStream_t stream = nullptr;
    int p_measure = P_MAXABS;   // pressure convergence measure (nim_set_p_measure)
    int device = 0;             // the device current at creation
    double *scratch = nullptr;  // this device's address of `scratch` (nim_end_step_noalloc)
    Dev dev;
    std::vector<int32_t> slot_ref, ref_slot;
    bool has_geo = false, has_pcoef = false;
    Ctl *h_ctl = nullptr;
    double *h_out = nullptr, *d_out = nullptr;
    std::vector<void *> allocs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gbatch = 0;
    int nred = 0;
    // sweep schedules (64- and 256-cell blocks) and their slab thickness
    int slab = 0;
    bool rslab_sweeps = true;  // sweeps: one radial slab per XCD (NIM_RSLAB=0: storage order)
    int tile = 0;              // FP64 source kernels from LDS tiles of tile x 256/tile cells (NIM_TILE=1; 0: per cell)
    // FP64 source kernels with two cells of a half-row per thread and 16-byte
    // loads (NIM_PAIR, default on where the shape allows): 5-12% faster per
    // kernel at 1M cells. The FP32 correction kernels stay one cell per
    // thread: paired, their occupancy fell from 3-8 to 2-3 waves per SIMD and
    // they ran 10-37% slower.
    bool pair = false;
    // sweeps load what a pass reads once (coefficients, metric differences,
    // FP32 residuals) non-temporally: from NT_MIN_CELLS cells on, where a
    // pass's arrays overflow the caches (1M cells: the FP32 pass 5% faster);
    // below, the arrays stay cached from pass to pass and plain loads are
    // faster (the mixed loop 16% at 110,592 cells). NIM_NT=0/1 forces it.
    bool nt = false;
    // FP64 sweeps: colour launches walking the planes alternately from the
    // bottom and from the top (NIM_KREV=0: always from the bottom), starting
    // from the bottom (sched256) or from the top (sched256r); Dev copy
    // dev_rev walks the per-cell kernels from the top
    std::vector<Sched> sched256, sched256r;
    Dev dev_rev;
    std::vector<Sched> dsched64, dsched256;  // FP32 correction sweeps (storage order: the radial slabs slow them)
    // FP64 loop: energy coefficients built inside the energy sweep instead
    // of stored by a kernel of their own (default; NIM_FUSE_T=0 turns it off)
    bool fuse_T = true;
    // mixed precision: FP32 corrections and constants, residual buffers
    int const_version = 0;        // bumped by every geometry or pressure-coefficient build
    struct {
        bool alloc = false;
        int version = -1;         // const_version the FP32 constants were cast from
        View<float> v32;          // corrections (sc, fc, src) and FP32 constants
        View<double> v64;         // FP64 state with the residual buffers
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        int gbatch = 0, ginner = 0;
    } mx;
    void *accel = nullptr;  // Anderson acceleration workspace (nim_accel.hip), made on first use
};

extern "C" const char *nim_last_error(void) { return g_err.c_str(); }

// for the other translation units of the library
int32_t nim_internal_fail(int32_t code, const char *msg) { return fail(code, msg); }
void *nim_internal_stream(nim_ctx *c) { return (void *)c->stream; }
const Dev *nim_internal_dev(nim_ctx *c) { return &c->dev; }
// the energy coefficients at the current state (the FP64 loop does not
// store them when fuse_T is set)
void nim_internal_energy_coefs(nim_ctx *c)
{
    hipLaunchKernelGGL(k_energy_coefs, dim3((c->N + BS - 1) / BS), dim3(BS), 0, c->stream, c->dev);
}

void nim_internal_lengths(nim_ctx *c, double *R_i, double *R_o)
{
    *R_i = c->prm.R_i;
    *R_o = c->prm.R_o;
}

static void init_schedules(nim_ctx *c);
void nim_internal_accel_free(void *ws);  // nim_accel.hip

// The device code addresses per-cell slot groups as slot * N + c and the
// face fields as field * NF + f in 32-bit integers. The largest products are
// the 45 pressure-coefficient slots (45 N) and the five face fields (5 NF),
// so both must stay within INT32_MAX: at most 47,721,858 cells.
static constexpr int64_t MAX_SLOTS_PER_CELL = 45;

static bool dims_ok(const nim_dims *d)
{
    // n_z >= 4: the four-point closures of the two adiabatic walls
    // (apply_zero_flux) would otherwise read each other's wall face
    if (!d || d->n_theta < 4 || d->n_theta % 2 != 0 || d->n_r < 3 || d->n_z < 4) return false;
    const int64_t NT = d->n_theta, NR = d->n_r, NZ = d->n_z;
    if (NT * NR > INT32_MAX) return false;  // keeps the next product within int64
    const int64_t N = NT * NR * NZ, NF = 3 * N + NT * NZ + NT * NR;
    return MAX_SLOTS_PER_CELL * N <= INT32_MAX && 5 * NF <= INT32_MAX;
}

extern "C" int32_t nim_counts(const nim_dims *d, int64_t *nmesh, int64_t *nface, int64_t *nbc)
{
    if (!dims_ok(d))
        return fail(NIM_E_ARG, "unsupported dimensions (n_theta even >= 4, n_r >= 3, n_z >= 4, "
                               "at most 47,721,858 cells)");
    int64_t NT = d->n_theta, NR = d->n_r, NZ = d->n_z;
    if (nmesh) *nmesh = NT * NR * NZ;
    if (nface) *nface = 3 * NT * NR * NZ + NT * NZ + NT * NR;  // parameters.f90:117
    if (nbc) *nbc = 2 * NT * NR + 2 * NT * NZ;
    return NIM_OK;
}

// reference face number (0-based) of each device face slot; the reference
// numbers faces by boundary group (create_geometry.f90:626-664): interior
// azimuthal faces, interior radial faces, interior axial faces, then
// bottom, top, hexagon and cylinder walls
extern "C" int32_t nim_face_map(const nim_dims *d, int32_t *out)
{
    int64_t N, NF, NBc;
    int32_t rc = nim_counts(d, &N, &NF, &NBc);
    if (rc) return rc;
    const int64_t NT = d->n_theta, NR = d->n_r, NZ = d->n_z;
    const int64_t o1 = N, o2 = o1 + NT * (NR - 1) * NZ, o3 = o2 + NT * NR * (NZ - 1), o4 = o3 + NT * NR;
    const int64_t o5 = o4 + NT * NR, o6 = o5 + NT * NZ;
    const int64_t oT = NT * (NR + 1) * NZ, oZ = oT + N;
    auto P = [&](int64_t i) { return (i & 1) * (NT / 2) + (i >> 1); };
    for (int64_t k = 0; k < NZ; k++)
        for (int64_t j = 0; j <= NR; j++)
            for (int64_t i = 0; i < NT; i++) {
                int64_t s = P(i) + NT * (j + (NR + 1) * k), r;
                if (j == 0) r = o5 + i + NT * k;
                else if (j == NR) r = o6 + i + NT * k;
                else r = o1 + i + NT * ((j - 1) + (NR - 1) * k);
                out[s] = (int32_t)r;
            }
    for (int64_t c = 0; c < N; c++) out[oT + P(c % NT) + (c / NT) * NT] = (int32_t)c;
    for (int64_t k = 0; k <= NZ; k++)
        for (int64_t j = 0; j < NR; j++)
            for (int64_t i = 0; i < NT; i++) {
                int64_t s = oZ + P(i) + NT * (j + NR * k), r;
                if (k == 0) r = o3 + i + NT * j;
                else if (k == NZ) r = o4 + i + NT * j;
                else r = o2 + i + NT * (j + NR * (k - 1));
                out[s] = (int32_t)r;
            }
    return NIM_OK;
}

template <class T>
static int32_t dalloc(nim_ctx *c, T **p, size_t n)
{
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 8));
    HIPCHK(hipMemset(q, 0, n * sizeof(T)));
    c->allocs.push_back(q);
    *p = (T *)q;
    return NIM_OK;
}

extern "C" int32_t nim_ctx_create(const nim_dims *dims, const nim_params *prm, nim_ctx **out)
{
    if (!out || !prm) return fail(NIM_E_ARG, "null argument");
    *out = nullptr;
    int64_t N, NF, NB;
    int32_t rc = nim_counts(dims, &N, &NF, &NB);
    if (rc) return rc;
    nim_ctx *c = new nim_ctx();
    c->dims = *dims;
    c->prm = *prm;
    c->NT = dims->n_theta; c->NR = dims->n_r; c->NZ = dims->n_z;
    c->N = (int)N; c->NF = (int)NF; c->NB = (int)NB;
    c->oT = c->NT * (c->NR + 1) * c->NZ;
    c->oZ = c->oT + c->N;
    c->slot_ref.resize(NF);
    nim_face_map(dims, c->slot_ref.data());
    c->ref_slot.resize(NF);
    for (int s = 0; s < c->NF; s++) c->ref_slot[c->slot_ref[s]] = s;
    Dev &d = c->dev;
    d.NT = c->NT; d.NR = c->NR; d.NZ = c->NZ; d.N = c->N; d.NF = c->NF; d.NB = c->NB; d.oT = c->oT; d.oZ = c->oZ;
    d.nu = prm->nu; d.Rinv = 1 / prm->Re; d.alpha = prm->alpha; d.rho = prm->rho; d.rinv = 1 / prm->rho;
    d.g_beta = prm->g_beta; d.DT = prm->DT;
    d.w_u = prm->w_u; d.w_v = prm->w_v; d.w_w = prm->w_w; d.w_p = prm->w_p; d.w_p2 = prm->w_p2; d.w_T = prm->w_T;
    d.v_limit = prm->v_limit;
    c->nred = (c->N + BS - 1) / BS;
    {
        const char *e = getenv("NIM_KFAST");
        d.kfast = e ? atoi(e) != 0 : 0;
        const char *r = getenv("NIM_RSLAB");
        c->rslab_sweeps = r ? atoi(r) != 0 : true;
        d.rslab = !d.kfast && c->rslab_sweeps && c->NR % 8 == 0 &&
                  ((size_t)c->NT * (c->NR / 8)) % BS == 0;
        const char *tl = getenv("NIM_TILE");
        const bool tiles = tl ? atoi(tl) != 0 : false;
        c->tile = !tiles ? 0 : c->NT % 64 == 0 ? 64 : c->NT % 32 == 0 ? 32 : 0;
        // two cells per thread need even half-rows and, with radial slabs,
        // slabs of whole 2*BS-cell blocks
        const char *pr = getenv("NIM_PAIR");
        const bool pairs = pr ? atoi(pr) != 0 : true;
        c->pair = pairs && !c->tile && !d.kfast && c->NT % 4 == 0 &&
                  (!d.rslab || ((size_t)c->NT * (c->NR / 8)) % (2 * BS) == 0);
        const char *f = getenv("NIM_FUSE_T");
        c->fuse_T = f ? atoi(f) != 0 : true;
        const char *n = getenv("NIM_NT");
        c->nt = n ? atoi(n) != 0 : c->N >= NT_MIN_CELLS;
    }
    init_schedules(c);
    double *gc, *An, *wall;
    int32_t *wpts;
#define TRY(x)                                                                                          \
    do {                                                                                                \
        int32_t r_ = (x);                                                                               \
        if (r_) { nim_ctx_destroy(c); return r_; }                                                      \
    } while (0)
    TRY(dalloc(c, &gc, (size_t)NGC * N));
    TRY(dalloc(c, &An, (size_t)3 * NF));
    TRY(dalloc(c, &d.AA, (size_t)45 * N));
    TRY(dalloc(c, &d.AQ, (size_t)3 * NAQ * N));
    TRY(dalloc(c, &wall, (size_t)NWS * NB));
    TRY(dalloc(c, &wpts, (size_t)5 * NB));
    TRY(dalloc(c, &d.sc, (size_t)NSC * N));
    TRY(dalloc(c, &d.fc, (size_t)5 * NF));
    TRY(dalloc(c, &d.MA, (size_t)18 * N));
    TRY(dalloc(c, &d.TA, (size_t)18 * N));
    TRY(dalloc(c, &d.MF, (size_t)NCF * N));
    TRY(dalloc(c, &d.TF, (size_t)NCF * N));
    TRY(dalloc(c, &d.src, (size_t)NSRC * N));
    TRY(dalloc(c, &d.red, (size_t)6 * c->nred + 64));
    TRY(dalloc(c, &d.ctl, (size_t)1));
    TRY(dalloc(c, &d.errs, (size_t)2 * NERR * ERR_COPIES * ERR_STRIDE));
    TRY(dalloc(c, &c->d_out, (size_t)16));
    d.gc = gc; d.An = An; d.wall = wall; d.wpts = wpts;
    {
        const char *kr = getenv("NIM_KREV");
        const bool krev = kr ? atoi(kr) != 0 : true;
        c->dev_rev = d;
        c->dev_rev.krev = krev && d.rslab;
        if (!krev)
            for (auto *v : {&c->sched256, &c->sched256r})
                for (Sched &sc : *v) sc.rev = 0;
    }
    if (hipHostMalloc((void **)&c->h_ctl, sizeof(Ctl), 0) != hipSuccess ||
        hipHostMalloc((void **)&c->h_out, 16 * sizeof(double), 0) != hipSuccess) {
        nim_ctx_destroy(c);
        return fail(NIM_E_HIP, "hipHostMalloc failed");
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        nim_ctx_destroy(c);
        return fail(NIM_E_HIP, "hipStreamCreate failed");
    }
    // set-up of the allocation-free entry point on this device (null if
    // unavailable: nim_end_step_noalloc then reports NIM_E_HIP)
    if (hipGetDevice(&c->device) != hipSuccess) {
        nim_ctx_destroy(c);
        return fail(NIM_E_HIP, "hipGetDevice failed");
    }
    c->scratch = noalloc_scratch_current_device();
#undef TRY
    *out = c;
    return NIM_OK;
}

extern "C" int32_t nim_ctx_destroy(nim_ctx *c)
{
    if (!c) return NIM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->mx.gexec) (void)hipGraphExecDestroy(c->mx.gexec);
    if (c->mx.graph) (void)hipGraphDestroy(c->mx.graph);
    nim_internal_accel_free(c->accel);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_out) (void)hipHostFree(c->h_out);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NIM_OK;
}

extern "C" void *nim_stream(nim_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int32_t nim_synchronize(nim_ctx *c)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return NIM_OK;
}

static int32_t h2d(nim_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NIM_OK;
}

static int32_t d2h(nim_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NIM_OK;
}

// wall face b -> (position, i, j, k) in the reference's boundary order
static void wall_of(const nim_ctx *c, int b, int &pos, int &i, int &j, int &k)
{
    const int NT = c->NT, NR = c->NR, NZ = c->NZ;
    if (b < NT * NR) { pos = 5; i = b % NT; j = b / NT; k = 0; }
    else if (b < 2 * NT * NR) { pos = 6; i = (b - NT * NR) % NT; j = (b - NT * NR) / NT; k = NZ - 1; }
    else if (b < 2 * NT * NR + NT * NZ) { pos = 1; i = (b - 2 * NT * NR) % NT; k = (b - 2 * NT * NR) / NT; j = 0; }
    else { pos = 2; i = (b - 2 * NT * NR - NT * NZ) % NT; k = (b - 2 * NT * NR - NT * NZ) / NT; j = NR - 1; }
}

// storage position of azimuthal index i (even i first, then odd i)
static inline int hpos(const nim_ctx *c, int i) { return (i & 1) * (c->NT / 2) + (i >> 1); }

// device slot of face `pos` of cell (i,j,k); must agree with fidx()
static int hslot(const nim_ctx *c, int pos, int i, int j, int k)
{
    const int NT = c->NT, NR = c->NR;
    switch (pos) {
    case 1: return hpos(c, i) + NT * (j + (NR + 1) * k);
    case 2: return hpos(c, i) + NT * (j + 1 + (NR + 1) * k);
    case 3: return c->oT + hpos(c, i) + NT * (j + NR * k);
    case 4: return c->oT + hpos(c, (i + 1) % NT) + NT * (j + NR * k);
    case 5: return c->oZ + hpos(c, i) + NT * (j + NR * k);
    default: return c->oZ + hpos(c, i) + NT * (j + NR * (k + 1));
    }
}

// device index of the reference's (natural-order) cell n
static inline int hcell(const nim_ctx *c, int n) { return hpos(c, n % c->NT) + (n / c->NT) * c->NT; }

extern "C" int32_t nim_upload_geometry(nim_ctx *c, const nim_geometry *g)
{
    if (!c || !g || !g->volume || !g->xi_x || !g->gij || !g->gij_faces || !g->G1 || !g->G2 || !g->G3 ||
        !g->faces_area || !g->faces_normal || !g->faces_mid || !g->bc_points || !g->bc_a)
        return fail(NIM_E_ARG, "nim_upload_geometry: missing array");
    const int N = c->N, NF = c->NF, NB = c->NB, NT = c->NT, NR = c->NR, NZ = c->NZ;
    std::vector<double> gc((size_t)NGC * N, 0.), An((size_t)3 * NF), wall((size_t)NWS * NB, 0.);
    std::vector<int32_t> wpts((size_t)5 * NB);
    auto GF = [&](int n, int a, int b, int f) { return g->gij_faces[(size_t)n * 54 + (a - 1) + 3 * (b - 1) + 9 * (f - 1)]; };
    auto GIJ = [&](int n, int a, int b) { return g->gij[(size_t)n * 9 + (a - 1) + 3 * (b - 1)]; };
    for (int n = 0; n < N; n++) {
        int i = n % NT, j = (n / NT) % NR, k = n / (NT * NR);
        const int dn = hcell(c, n);
        gc[(size_t)G_INVV * N + dn] = 1. / g->volume[n];
        for (int a = 1; a <= 3; a++)
            for (int b = 1; b <= 3; b++)
                gc[(size_t)(G_XI + 3 * (a - 1) + (b - 1)) * N + dn] = g->xi_x[(size_t)n * 9 + (a - 1) + 3 * (b - 1)];
        for (int a = 1; a <= 3; a++) gc[(size_t)(G_GD + a - 1) * N + dn] = GIJ(n, a, a);
        const int go[6][2] = {{1, 2}, {2, 1}, {1, 3}, {3, 1}, {2, 3}, {3, 2}};
        for (int q = 0; q < 6; q++) gc[(size_t)(G_GO + q) * N + dn] = GIJ(n, go[q][0], go[q][1]);
        gc[(size_t)(G_G + 0) * N + dn] = g->G1[n];
        gc[(size_t)(G_G + 1) * N + dn] = g->G2[n];
        gc[(size_t)(G_G + 2) * N + dn] = g->G3[n];
        for (int p = 1; p <= 6; p++) gc[(size_t)(G_GF + p - 1) * N + dn] = GF(n, (p + 1) / 2, (p + 1) / 2, p);
        double *df = &gc[(size_t)G_DIFF * N];
        if (j < NR - 1) {
            int n2 = n + NT;
            df[0 * (size_t)N + dn] = GF(n2, 1, 2, 1) - GF(n, 1, 2, 2);
            df[1 * (size_t)N + dn] = GF(n2, 1, 3, 1) - GF(n, 1, 3, 2);
        }
        {
            int n2 = (i + 1) % NT + NT * (j + NR * k);
            df[2 * (size_t)N + dn] = GF(n2, 2, 1, 3) - GF(n, 2, 1, 4);
            df[3 * (size_t)N + dn] = GF(n2, 2, 3, 3) - GF(n, 2, 3, 4);
        }
        if (k < NZ - 1) {
            int n2 = n + NT * NR;
            df[4 * (size_t)N + dn] = GF(n2, 3, 1, 5) - GF(n, 3, 1, 6);
            df[5 * (size_t)N + dn] = GF(n2, 3, 2, 5) - GF(n, 3, 2, 6);
        }
    }
    for (int s = 0; s < NF; s++) {
        int r = c->slot_ref[s];
        for (int q = 0; q < 3; q++) An[(size_t)q * NF + s] = g->faces_area[r] * g->faces_normal[(size_t)r * 3 + q];
    }
    auto mid = [&](int r, int q) { return g->faces_mid[(size_t)r * 3 + q]; };
    auto dist = [&](int r1, int r2) {
        double dx = mid(r1, 0) - mid(r2, 0), dy = mid(r1, 1) - mid(r2, 1), dz = mid(r1, 2) - mid(r2, 2);
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    for (int b = 0; b < NB; b++) {
        int pos, i, j, k;
        wall_of(c, b, pos, i, j, k);
        int n = i + NT * (j + NR * k);
        int rw = c->slot_ref[hslot(c, pos, i, j, k)];
        for (int m = 0; m < 5; m++) {
            wall[(size_t)(W_A + m) * NB + b] = g->bc_a[(size_t)b * 20 + m];
            int rf = g->bc_points[(size_t)b * 5 + m] - 1;
            if (rf < 0 || rf >= NF) return fail(NIM_E_ARG, "bc_points out of range");
            wpts[(size_t)m * NB + b] = c->ref_slot[rf];
        }
        for (int q = 0; q < 3; q++) wall[(size_t)(W_N + q) * NB + b] = g->faces_normal[(size_t)rw * 3 + q];
        double gB2, gB2c, gB3;
        if (pos <= 2) { gB2 = gB2c = GF(n, 1, 2, pos); gB3 = GF(n, 1, 3, pos); }
        else {
            gB2 = GF(n, 3, 1, pos);
            gB2c = pos == 6 ? GF(n, 3, 2, pos) : gB2;  // bc.f90:318
            gB3 = GF(n, 3, 2, pos);
        }
        wall[(size_t)W_GB2 * NB + b] = gB2;
        wall[(size_t)W_GB2C * NB + b] = gB2c;
        wall[(size_t)W_GB3 * NB + b] = gB3;
        // one-sided 4-point weights from face mid-points (bc.f90:353-387,
        // nse.f90:2722-2782): the wall face and the three next parallel faces
        int r2, r3, r4;
        if (pos == 5) { r2 = hslot(c, 6, i, j, 0); r3 = hslot(c, 6, i, j, 1); r4 = hslot(c, 6, i, j, 2); }
        else if (pos == 6) { r2 = hslot(c, 5, i, j, k); r3 = hslot(c, 5, i, j, k - 1); r4 = hslot(c, 5, i, j, k - 2); }
        else if (pos == 1) { r2 = hslot(c, 2, i, 0, k); r3 = hslot(c, 2, i, 1, k); r4 = hslot(c, 2, i, 2, k); }
        else { r2 = hslot(c, 1, i, j, k); r3 = hslot(c, 1, i, j - 1, k); r4 = hslot(c, 1, i, j - 2, k); }
        r2 = c->slot_ref[r2]; r3 = c->slot_ref[r3]; r4 = c->slot_ref[r4];
        double h1 = dist(rw, r2), h2 = dist(r2, r3), h3 = dist(r3, r4);
        double A = ((h1 + h2) * (h1 + h2 + h3)) / (h1 * h2 * (h2 + h3));
        double Bw = (h1 * (h1 + h2 + h3)) / (h2 * h3 * (h1 + h2));
        double Cw = (h1 * (h1 + h2)) / (h3 * (h2 + h3) * (h1 + h2 + h3));
        double S = (1. / h1 + 1. / (h1 + h2) + 1. / (h1 + h2 + h3));
        if (pos >= 5) {
            wall[(size_t)(W_ZF + 0) * NB + b] = A;
            wall[(size_t)(W_ZF + 1) * NB + b] = Bw;
            wall[(size_t)(W_ZF + 2) * NB + b] = Cw;
            wall[(size_t)(W_ZF + 3) * NB + b] = S;
        } else {
            wall[(size_t)(W_NU + 0) * NB + b] = S;
            wall[(size_t)(W_NU + 1) * NB + b] = A;
            wall[(size_t)(W_NU + 2) * NB + b] = Bw;
            wall[(size_t)(W_NU + 3) * NB + b] = Cw;
            wall[(size_t)(W_NU + 4) * NB + b] = g->faces_area[rw];
        }
    }
    int32_t rc;
    if ((rc = h2d(c, (void *)c->dev.gc, gc.data(), gc.size() * 8))) return rc;
    if ((rc = h2d(c, (void *)c->dev.An, An.data(), An.size() * 8))) return rc;
    // keep BCs_faces_ap slots if they were already built on the device
    if (c->has_pcoef) {
        std::vector<double> ap((size_t)8 * NB);
        if ((rc = d2h(c, ap.data(), c->dev.wall + (size_t)W_AP * NB, ap.size() * 8))) return rc;
        std::memcpy(&wall[(size_t)W_AP * NB], ap.data(), ap.size() * 8);
    }
    if ((rc = h2d(c, c->dev.wall, wall.data(), wall.size() * 8))) return rc;
    if ((rc = h2d(c, (void *)c->dev.wpts, wpts.data(), wpts.size() * 4))) return rc;
    c->has_geo = true;
    c->const_version++;
    return NIM_OK;
}

static inline int nblk(int n) { return (n + BS - 1) / BS; }

extern "C" int32_t nim_build_pressure_coeffs(nim_ctx *c)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    if (!c->has_geo) return fail(NIM_E_STATE, "geometry not uploaded");
    hipLaunchKernelGGL(k_pressure_coeffs, dim3(nblk(c->N)), dim3(BS), 0, c->stream, c->dev);
    hipLaunchKernelGGL(k_wall_pressure_coeffs, dim3(nblk(c->NB)), dim3(BS), 0, c->stream, c->dev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_pcoef = true;
    c->const_version++;
    return NIM_OK;
}

// ---- named arrays in reference layout -------------------------------
struct ArrRef {
    double *dev;   // first element
    int kind;      // 0 cells (stride N), 1 faces (ref numbering), 2 walls (stride NB)
    int mult;      // inner (Fortran) dimension
};

static bool lookup(nim_ctx *c, const char *name, ArrRef &a)
{
    const int N = c->N;
    Dev &d = c->dev;
    std::string s(name);
    const char *ff[5] = {"faces_u", "faces_v", "faces_w", "faces_p", "faces_T"};
    for (int q = 0; q < 5; q++)
        if (s == ff[q]) { a = {d.fc + (size_t)q * c->NF, 1, 1}; return true; }
    const char *cf[4] = {"u", "v", "w", "T"};
    for (int q = 0; q < 4; q++) {
        if (s == std::string(cf[q]) + "xyz") { a = {d.sc + (size_t)(S_CUR + q) * N, 0, 1}; return true; }
        if (s == std::string(cf[q]) + "xyz_old") { a = {d.sc + (size_t)(S_OLD + q) * N, 0, 1}; return true; }
        if (s == std::string(cf[q]) + "xyzt") { a = {d.sc + (size_t)(S_TC + q) * N, 0, 1}; return true; }
    }
    if (s == "meshes_p") { a = {d.sc + (size_t)S_P * N, 0, 1}; return true; }
    if (s == "meshes_dila") { a = {d.sc + (size_t)S_DILA * N, 0, 1}; return true; }
    if (s == "M_F4") { a = {d.MF + (size_t)C_F4 * N, 0, 8}; return true; }
    if (s == "M_F1") { a = {d.MF + (size_t)C_F1 * N, 0, 5}; return true; }
    if (s == "M_F2") { a = {d.MF + (size_t)C_F2 * N, 0, 5}; return true; }
    if (s == "M_F3") { a = {d.MF + (size_t)C_F3 * N, 0, 5}; return true; }
    if (s == "MT_F4") { a = {d.TF + (size_t)C_F4 * N, 0, 8}; return true; }
    if (s == "MT_F1") { a = {d.TF + (size_t)C_F1 * N, 0, 5}; return true; }
    if (s == "MT_F2") { a = {d.TF + (size_t)C_F2 * N, 0, 5}; return true; }
    if (s == "MT_F3") { a = {d.TF + (size_t)C_F3 * N, 0, 5}; return true; }
    if (s == "AA_px") { a = {d.AA, 0, 15}; return true; }
    if (s == "AA_py") { a = {d.AA + (size_t)15 * N, 0, 15}; return true; }
    if (s == "AA_pz") { a = {d.AA + (size_t)30 * N, 0, 15}; return true; }
    if (s == "bc_ap") { a = {d.wall + (size_t)W_AP * c->NB, 2, 8}; return true; }
    if (s == "R_B_u") { a = {d.src + (size_t)R_U * N, 0, 1}; return true; }
    if (s == "R_B_v") { a = {d.src + (size_t)R_V * N, 0, 1}; return true; }
    if (s == "R_B_w") { a = {d.src + (size_t)R_W * N, 0, 1}; return true; }
    if (s == "R_b_p") { a = {d.src + (size_t)R_P * N, 0, 1}; return true; }
    if (s == "R_D_T") { a = {d.src + (size_t)R_T * N, 0, 1}; return true; }
    return false;
}

static int32_t move_ref(nim_ctx *c, const ArrRef &a, const char *name, double *host, int64_t count, bool to_host);

static int32_t move_array(nim_ctx *c, const char *name, double *host, int64_t count, bool to_host)
{
    if (!c || !name || !host) return fail(NIM_E_ARG, "null argument");
    ArrRef a;
    if (!lookup(c, name, a)) return fail(NIM_E_ARG, std::string("unknown array ") + name);
    return move_ref(c, a, name, host, count, to_host);
}

// a device array of the library's layout (kind 0 cells [mult][N], 1 faces,
// 2 walls [mult][NB]) to or from the reference layout, for the other
// translation units
int32_t nim_internal_move(nim_ctx *c, double *dev, int kind, int mult, const char *name, double *host, int64_t count,
                          bool to_host)
{
    return move_ref(c, ArrRef{dev, kind, mult}, name, host, count, to_host);
}

static int32_t move_ref(nim_ctx *c, const ArrRef &a, const char *name, double *host, int64_t count, bool to_host)
{
    const int rows = a.kind == 1 ? c->NF : a.kind == 2 ? c->NB : c->N;
    if (count != (int64_t)rows * a.mult) return fail(NIM_E_ARG, std::string("wrong element count for ") + name);
    std::vector<double> buf((size_t)rows * a.mult);
    int32_t rc;
    if (to_host) {
        if ((rc = d2h(c, buf.data(), a.dev, buf.size() * 8))) return rc;
        if (a.kind == 1) {
            for (int s = 0; s < rows; s++) host[c->slot_ref[s]] = buf[s];
        } else {
            for (int r = 0; r < rows; r++) {
                int dr = a.kind == 0 ? hcell(c, r) : r;
                for (int m = 0; m < a.mult; m++) host[(size_t)r * a.mult + m] = buf[(size_t)m * rows + dr];
            }
        }
        return NIM_OK;
    }
    if (a.kind == 1) {
        for (int s = 0; s < rows; s++) buf[s] = host[c->slot_ref[s]];
    } else {
        for (int r = 0; r < rows; r++) {
            int dr = a.kind == 0 ? hcell(c, r) : r;
            for (int m = 0; m < a.mult; m++) buf[(size_t)m * rows + dr] = host[(size_t)r * a.mult + m];
        }
    }
    return h2d(c, a.dev, buf.data(), buf.size() * 8);
}

extern "C" int32_t nim_get_array(nim_ctx *c, const char *name, double *out, int64_t count)
{
    return move_array(c, name, out, count, true);
}

extern "C" int32_t nim_set_array(nim_ctx *c, const char *name, const double *in, int64_t count)
{
    return move_array(c, name, const_cast<double *>(in), count, false);
}

static int32_t state_io(nim_ctx *c, const nim_state *st, bool to_host)
{
    if (!c || !st) return fail(NIM_E_ARG, "null argument");
    struct { const char *n; double *p; int64_t cnt; } tab[] = {
        {"faces_u", st->faces_u, c->NF}, {"faces_v", st->faces_v, c->NF}, {"faces_w", st->faces_w, c->NF},
        {"faces_p", st->faces_p, c->NF}, {"faces_T", st->faces_T, c->NF},
        {"uxyz", st->uxyz, c->N}, {"vxyz", st->vxyz, c->N}, {"wxyz", st->wxyz, c->N}, {"Txyz", st->Txyz, c->N},
        {"uxyz_old", st->uxyz_old, c->N}, {"vxyz_old", st->vxyz_old, c->N}, {"wxyz_old", st->wxyz_old, c->N},
        {"Txyz_old", st->Txyz_old, c->N}, {"uxyzt", st->uxyzt, c->N}, {"vxyzt", st->vxyzt, c->N},
        {"wxyzt", st->wxyzt, c->N}, {"Txyzt", st->Txyzt, c->N}, {"meshes_p", st->meshes_p, c->N},
        {"meshes_dila", st->meshes_dila, c->N}};
    for (auto &t : tab) {
        if (!t.p) continue;
        int32_t rc = move_array(c, t.n, t.p, t.cnt, to_host);
        if (rc) return rc;
    }
    return NIM_OK;
}

extern "C" int32_t nim_upload_state(nim_ctx *c, const nim_state *st) { return state_io(c, st, false); }
extern "C" int32_t nim_download_state(nim_ctx *c, nim_state *st) { return state_io(c, st, true); }

extern "C" int32_t nim_build_momentum_coeffs(nim_ctx *c)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    if (!c->has_geo) return fail(NIM_E_STATE, "geometry not uploaded");
    hipLaunchKernelGGL(k_coef_A, dim3(nblk(c->N)), dim3(BS), 0, c->stream, c->dev, 0);
    hipLaunchKernelGGL(k_coef_F, dim3(nblk(c->N)), dim3(BS), 0, c->stream, c->dev, 0);
    HIPCHK(hipGetLastError());
    return NIM_OK;
}

// ---- one outer iteration ---------------------------------------------
// momentum kernels: 64 cells per block, one wave per velocity component
static const dim3 BS3(64, 3);
// blocks of the per-cell kernels (storage order or plane-fastest)
static int cell_grid(const nim_ctx *c, int bs)
{
    if (!c->dev.kfast) return (c->N + bs - 1) / bs;
    return ((c->NT * c->NR + bs - 1) / bs) * c->NZ;
}
static int cell_blocks3(const nim_ctx *c) { return cell_grid(c, 64); }

// blocks of a two-cells-per-thread source launch
static int pair_grid(const nim_ctx *c) { return (c->N + 2 * BS - 1) / (2 * BS); }

// blocks of a tiled source launch: one per tile of a plane
static int tile_blocks(const nim_ctx *c)
{
    const int TJ = 256 / c->tile;
    return (c->NT / c->tile) * ((c->NR + TJ - 1) / TJ) * c->NZ;
}

// energy sources (update_temperature_sources): from LDS tiles where the
// azimuthal count allows, else one thread per cell
static void launch_T_src(const nim_ctx *c, hipStream_t s)
{
    const Dev &d = c->dev;
    if (c->tile) {
        const int TJ = 256 / c->tile, blocks = (c->NT / c->tile) * ((c->NR + TJ - 1) / TJ) * c->NZ;
        if (c->tile == 64) hipLaunchKernelGGL(k_T_src_tile<64>, dim3(blocks), dim3(256), 0, s, d);
        else hipLaunchKernelGGL(k_T_src_tile<32>, dim3(blocks), dim3(256), 0, s, d);
    } else if (c->pair) {
        hipLaunchKernelGGL(k_T_src2, dim3(pair_grid(c)), dim3(BS), 0, s, d);
    } else {
        hipLaunchKernelGGL(k_T_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d);
    }
}

// Launches of one relaxation sweep. With slabs of `slab` axial planes
// (S slabs), launch L runs colour c on slab s for every s + 2c = L. Colour
// c on slab s then starts after colour c - 1 finished on slabs s - 1, s,
// s + 1 and before any higher colour touched them, and no stencil (reach 2
// planes at the walls, so slab >= 2) crosses more than one slab: every cell
// reads exactly what it reads in the four full-domain colour phases, which
// are the one-slab case. Colour c + 1 comes back to a slab two launches
// later, while its data are still in the caches.
static std::vector<Sched> build_schedule(const nim_ctx *c, int cells_per_block, bool rslab, int first_rev = 0)
{
    const int NZ = c->NZ, ks = c->slab, S = (NZ + ks - 1) / ks;
    const int per_plane = (c->NT / 2) * ((c->NR + 1) / 2);
    std::vector<Sched> out;
    for (int L = 0; L < S + 6; L++) {
        Sched sc{};
        int blocks = 0;
        for (int col = 0; col < 4; col++) {
            int sl = L - 2 * col;
            if (sl < 0 || sl >= S) continue;
            int q = sc.n++;
            sc.color[q] = col;
            sc.k0[q] = sl * ks;
            sc.k1[q] = std::min(NZ, (sl + 1) * ks);
            sc.bstart[q] = blocks;
            blocks += (per_plane * (sc.k1[q] - sc.k0[q]) + cells_per_block - 1) / cells_per_block;
        }
        if (!sc.n) continue;
        sc.bstart[sc.n] = blocks;
        // one item over all planes: one radial slab per XCD when the
        // colour's rows split evenly and in whole blocks
        const int nth = c->NT / 2, nrh = (c->NR + 1) / 2;
        if (rslab && sc.n == 1 && sc.k0[0] == 0 && sc.k1[0] == NZ && nrh % 8 == 0 &&
            (nth * (nrh / 8)) % cells_per_block == 0)
            sc.rsb = nth * (nrh / 8) / cells_per_block;
        // a full-domain colour launch on radial slabs walks the planes in the
        // direction opposite to the launch before it
        sc.rev = sc.rsb ? (first_rev + (int)out.size()) & 1 : 0;
        out.push_back(sc);
    }
    return out;
}

static void init_schedules(nim_ctx *c)
{
    const char *e = getenv("NIM_SLAB");
    int ks = e ? atoi(e) : 0;
    c->slab = ks >= 2 && ks < c->NZ ? ks : c->NZ;
    c->sched256 = build_schedule(c, BS, c->rslab_sweeps, 0);
    c->sched256r = build_schedule(c, BS, c->rslab_sweeps, 1);
    c->dsched64 = build_schedule(c, 64, false);
    c->dsched256 = build_schedule(c, BS, false);
}

static inline int sched_blocks(const Sched &s) { return s.bstart[s.n]; }

// `mark` is called after every launch with the kernel class (timing passes).
// With fused (the FP64 loop when fuse_T is set), stage 4 only updates the
// adiabatic walls and stage 6 builds the energy coefficients itself.
template <class Mark>
static void launch_stage(nim_ctx *c, int stage, Mark mark, bool fused = false)
{
    const Dev &d = c->dev;
    hipStream_t s = c->stream;
    const int gS = cell_grid(c, BS);
    switch (stage) {
    // plane directions along a pass: momentum sources up, momentum sweep
    // down-up-down-up, pressure sources down, pressure sweep up-down-up-down,
    // energy sources up, energy sweep down-up-down-up (with radial slabs)
    case 0:
        if (c->tile == 64) hipLaunchKernelGGL(k_mom_src_tile<64>, dim3(tile_blocks(c)), dim3(BS), 0, s, d);
        else if (c->tile == 32) hipLaunchKernelGGL(k_mom_src_tile<32>, dim3(tile_blocks(c)), dim3(BS), 0, s, d);
        else if (c->pair) hipLaunchKernelGGL(k_mom_src2, dim3(pair_grid(c)), dim3(BS), 0, s, d);
        else hipLaunchKernelGGL(k_mom_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d);
        mark(0);
        break;
    case 1:
        for (const Sched &sc : c->sched256r) {
            hipLaunchKernelGGL(c->nt ? k_mom_sweep<true> : k_mom_sweep<false>, dim3(sched_blocks(sc)), dim3(BS), 0,
                               s, d, sc);
            mark(1);
        }
        break;
    case 2:
        if (c->tile == 64) hipLaunchKernelGGL(k_ppe_src_tile<64>, dim3(tile_blocks(c)), dim3(BS), 0, s, d);
        else if (c->tile == 32) hipLaunchKernelGGL(k_ppe_src_tile<32>, dim3(tile_blocks(c)), dim3(BS), 0, s, d);
        else if (c->pair) hipLaunchKernelGGL(k_ppe_src2, dim3(pair_grid(c)), dim3(BS), 0, s, c->dev_rev);
        else hipLaunchKernelGGL(k_ppe_src, dim3(gS), dim3(BS), 0, s, c->dev_rev);
        mark(2);
        break;
    case 3:
        for (const Sched &sc : c->sched256) {
            hipLaunchKernelGGL(c->nt ? k_ppe_sweep<true> : k_ppe_sweep<false>, dim3(sched_blocks(sc)), dim3(BS), 0,
                               s, d, sc);
            mark(3);
        }
        break;
    case 4:
        if (fused)
            hipLaunchKernelGGL(k_zero_flux, dim3((2 * c->NT * c->NR + BS - 1) / BS), dim3(BS), 0, s, d);
        else
            hipLaunchKernelGGL(k_T_coef, dim3(gS), dim3(BS), 0, s, d);
        mark(4);
        break;
    case 5: launch_T_src(c, s); mark(6); break;
    case 6:
        for (const Sched &sc : c->sched256r) {
            if (fused)
                hipLaunchKernelGGL(c->nt ? k_T_sweep_fused<true> : k_T_sweep_fused<false>, dim3(sched_blocks(sc)),
                                   dim3(BS), 0, s, d, sc);
            else hipLaunchKernelGGL(k_T_sweep, dim3(sched_blocks(sc)), dim3(BS), 0, s, d, sc);
            mark(7);
        }
        break;
    }
}

static void launch_stage(nim_ctx *c, int stage)
{
    launch_stage(c, stage, [](int) {});
}

static void launch_iteration(nim_ctx *c)
{
    for (int st = 0; st <= 6; st++) launch_stage(c, st, [](int) {}, c->fuse_T);
    hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, c->stream, c->dev);
}

// the pressure measure of a solve: the context's (nim_set_p_measure); in the
// mixed loops the half range is taken of each cycle's pressure update, since
// the FP64 residual of the drifting rest state keeps a range (nim_kernels.h)
static int p_measure_of(const nim_ctx *c, bool mixed)
{
    return mixed && c->p_measure == P_HALFRANGE ? (int)P_HALFRANGE_UPDATE : c->p_measure;
}

static int32_t reset_ctl(nim_ctx *c, const double tol[5], int max_iter, int min_iter, bool mixed = false)
{
    hipLaunchKernelGGL(k_solve_reset, dim3(1), dim3(256), 0, c->stream, c->dev, tol[0], tol[1], tol[2], tol[3],
                       tol[4], max_iter, min_iter, p_measure_of(c, mixed));
    HIPCHK(hipGetLastError());
    return NIM_OK;
}

extern "C" int32_t nim_set_p_measure(nim_ctx *c, int32_t mode)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    if (mode != NIM_P_MAXABS && mode != NIM_P_HALFRANGE) return fail(NIM_E_ARG, "unknown pressure measure");
    c->p_measure = mode;
    return NIM_OK;
}

extern "C" int32_t nim_run_stage(nim_ctx *c, int32_t stage)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    if (stage < 0 || stage > 6) return fail(NIM_E_ARG, "stage out of range");
    const double tol[5] = {0, 0, 0, 0, 0};
    int32_t rc = reset_ctl(c, tol, 1 << 30, 0);
    if (rc) return rc;
    launch_stage(c, stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return NIM_OK;
}

static int batch_size()
{
    const char *e = getenv("NIM_BATCH");
    int b = e ? atoi(e) : 16;
    return b < 1 ? 1 : b > 256 ? 256 : b;
}

static bool use_graph()
{
    const char *e = getenv("NIM_GRAPH");
    return !(e && e[0] == '0');
}

static int32_t launch_batch(nim_ctx *c, int batch)
{
    if (!use_graph()) {
        for (int q = 0; q < batch; q++) launch_iteration(c);
        HIPCHK(hipGetLastError());
        return NIM_OK;
    }
    if (!c->gexec || c->gbatch != batch) {
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < batch; q++) launch_iteration(c);
        HIPCHK(hipStreamEndCapture(c->stream, &c->graph));
        HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->gbatch = batch;
    }
    HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    return NIM_OK;
}

extern "C" int32_t nim_inner_solve(nim_ctx *c, const double tol[5], int32_t max_iter, int32_t min_iter,
                                   int32_t *iters, double err[5])
{
    if (!c || !tol) return fail(NIM_E_ARG, "null argument");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    if (max_iter < 1) return fail(NIM_E_ARG, "max_iter must be >= 1");
    int32_t rc = reset_ctl(c, tol, max_iter, min_iter);
    if (rc) return rc;
    int batch = batch_size();
    for (;;) {
        if ((rc = launch_batch(c, batch))) return rc;
        HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->h_ctl->done) break;
    }
    if (iters) *iters = c->h_ctl->iter;
    if (err)
        for (int q = 0; q < 5; q++) err[q] = c->h_ctl->err[q];
    int st = c->h_ctl->status;
    if (st == 3) return fail(NIM_E_MAXITER, "inner iteration cap reached");
    if (st == 4) return fail(NIM_E_NONFINITE, "non-finite increment");
    return NIM_OK;
}

// ---- mixed precision ----------------------------------------------------
// FP32 buffers of the correction form (allocated on first use) and FP32
// copies of the constant arrays, recast when geometry or pressure
// coefficients were rebuilt
static int32_t mixed_prepare(nim_ctx *c)
{
    auto &m = c->mx;
    const size_t N = c->N, NF = c->NF, NB = c->NB;
    if (!m.alloc) {
        float *gc, *An, *AQ, *wall, *MF, *TF, *sc, *fc, *src, *rc, *rf, *grad;
        int32_t rc_;
        if ((rc_ = dalloc(c, &gc, NGC * N)) || (rc_ = dalloc(c, &An, 3 * NF)) || (rc_ = dalloc(c, &AQ, 3 * NAQ * N)) ||
            (rc_ = dalloc(c, &wall, NWS * NB)) || (rc_ = dalloc(c, &MF, NCF * N)) || (rc_ = dalloc(c, &TF, NCF * N)) ||
            (rc_ = dalloc(c, &sc, 4 * N)) || (rc_ = dalloc(c, &fc, 5 * NF)) || (rc_ = dalloc(c, &src, NSRC * N)) ||
            (rc_ = dalloc(c, &rc, 4 * N)) || (rc_ = dalloc(c, &rf, 5 * NF)) || (rc_ = dalloc(c, &grad, 9 * N)))
            return rc_;
        View<float> &v = m.v32;
        v.gc = gc; v.An = An; v.AQ = AQ; v.wall = wall; v.MF = MF; v.TF = TF;
        v.sc = sc; v.old = nullptr; v.fc = fc; v.src = src; v.rc = rc; v.rf = rf; v.TF32 = nullptr; v.grad = grad;
        m.v64 = view64(c->dev);
        m.v64.rc = rc; m.v64.rf = rf; m.v64.TF32 = TF; m.v64.grad = grad;
        m.alloc = true;
    }
    hipStream_t s = c->stream;
    auto cast = [&](const double *in, const float *out, size_t n) {
        int g = (int)std::min<size_t>((n + BS - 1) / BS, 4096);
        hipLaunchKernelGGL(k_cast, dim3(g), dim3(BS), 0, s, in, (float *)out, (long)n);
    };
    if (m.version != c->const_version) {
        const Dev &d = c->dev;
        cast(d.gc, m.v32.gc, NGC * N);
        cast(d.An, m.v32.An, 3 * NF);
        cast(d.AQ, m.v32.AQ, 3 * NAQ * N);
        cast(d.wall, m.v32.wall, NWS * NB);
        m.version = c->const_version;
    }
    cast(c->dev.MF, m.v32.MF, NCF * N);  // momentum coefficients change every step
    HIPCHK(hipGetLastError());
    return NIM_OK;
}

// FP64 residual of every unknown's equation at the current state, the
// maximum of |w r / a| per field reduced into the iteration control
template <class Mark>
static void launch_residual(nim_ctx *c, Mark mark)
{
    const Dev &d = c->dev;
    const View<double> &r = c->mx.v64;
    hipStream_t s = c->stream;
    const int gN = nblk(c->N);
    if (c->pair) hipLaunchKernelGGL(k_mom_src2, dim3(pair_grid(c)), dim3(BS), 0, s, d);
    else hipLaunchKernelGGL(k_mom_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d);
    mark(0);
    hipLaunchKernelGGL(k_res_mom, dim3(cell_blocks3(c)), BS3, 0, s, d, r); mark(0);
    if (c->pair) hipLaunchKernelGGL(k_res_ppe_src2, dim3(pair_grid(c)), dim3(BS), 0, s, d, r);
    else hipLaunchKernelGGL(k_res_ppe_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d, r);
    mark(1);
    hipLaunchKernelGGL(k_res_ppe, dim3(gN), dim3(BS), 0, s, d, r); mark(1);
    hipLaunchKernelGGL(k_res_T_coef, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d, r); mark(2);
    launch_T_src(c, s); mark(2);
    hipLaunchKernelGGL(k_res_T, dim3(gN), dim3(BS), 0, s, d, r); mark(2);
    hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, s, d); mark(3);
}

// one FP32 pass of the correction form, stages in the FP64 pass's order
template <class Mark>
static void launch_delta(nim_ctx *c, Mark mark)
{
    const Dev &d = c->dev;
    const View<float> &v = c->mx.v32;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_d_mom_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d, v); mark(4);
    for (const Sched &sc : c->dsched64) {
        hipLaunchKernelGGL(c->nt ? k_d_mom_sweep<true> : k_d_mom_sweep<false>, dim3(sched_blocks(sc)), BS3, 0, s, d,
                           v, sc);
        mark(5);
    }
    hipLaunchKernelGGL(k_d_ppe_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d, v); mark(6);
    for (const Sched &sc : c->dsched256) {
        hipLaunchKernelGGL(c->nt ? k_d_ppe_sweep<true> : k_d_ppe_sweep<false>, dim3(sched_blocks(sc)), dim3(BS), 0,
                           s, d, v, sc);
        mark(7);
    }
    hipLaunchKernelGGL(k_d_zero_flux, dim3(nblk(2 * c->NT * c->NR)), dim3(BS), 0, s, d, v); mark(8);
    hipLaunchKernelGGL(k_d_T_src, dim3(cell_grid(c, BS)), dim3(BS), 0, s, d, v); mark(9);
    for (const Sched &sc : c->dsched256) {
        hipLaunchKernelGGL(c->nt ? k_d_T_sweep<true> : k_d_T_sweep<false>, dim3(sched_blocks(sc)), dim3(BS), 0, s,
                           d, v, sc);
        mark(10);
    }
}

static void launch_residual(nim_ctx *c) { launch_residual(c, [](int) {}); }
static void launch_delta(nim_ctx *c) { launch_delta(c, [](int) {}); }

static void launch_update(nim_ctx *c, int inner)
{
    hipLaunchKernelGGL(k_mix_update, dim3(nblk(std::max(c->N, c->NF))), dim3(BS), 0, c->stream, c->dev, c->mx.v64,
                       c->mx.v32, 1. / inner);
}

static void launch_cycle(nim_ctx *c, int inner)
{
    launch_residual(c);
    for (int q = 0; q < inner; q++) launch_delta(c);
    launch_update(c, inner);
}

static int32_t launch_mixed_batch(nim_ctx *c, int batch, int inner)
{
    auto &m = c->mx;
    if (!use_graph()) {
        for (int q = 0; q < batch; q++) launch_cycle(c, inner);
        HIPCHK(hipGetLastError());
        return NIM_OK;
    }
    if (!m.gexec || m.gbatch != batch || m.ginner != inner) {
        if (m.gexec) { (void)hipGraphExecDestroy(m.gexec); m.gexec = nullptr; }
        if (m.graph) { (void)hipGraphDestroy(m.graph); m.graph = nullptr; }
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < batch; q++) launch_cycle(c, inner);
        HIPCHK(hipStreamEndCapture(c->stream, &m.graph));
        HIPCHK(hipGraphInstantiate(&m.gexec, m.graph, nullptr, nullptr, 0));
        m.gbatch = batch;
        m.ginner = inner;
    }
    HIPCHK(hipGraphLaunch(m.gexec, c->stream));
    return NIM_OK;
}

extern "C" int32_t nim_inner_solve_mixed(nim_ctx *c, const double tol[5], int32_t max_cycles, int32_t min_cycles,
                                         int32_t inner, int32_t *cycles, double err[5])
{
    if (!c || !tol) return fail(NIM_E_ARG, "null argument");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    if (max_cycles < 1) return fail(NIM_E_ARG, "max_cycles must be >= 1");
    if (inner < 1 || inner > 1024) return fail(NIM_E_ARG, "inner passes must be in [1, 1024]");
    int32_t rc = mixed_prepare(c);
    if (rc) return rc;
    if ((rc = reset_ctl(c, tol, max_cycles, min_cycles, true))) return rc;
    // cycles between host reads of the control block: about 16 passes
    const char *e = getenv("NIM_MIXED_BATCH");
    int batch = e ? atoi(e) : (16 + inner) / (inner + 1);
    batch = batch < 1 ? 1 : batch > 64 ? 64 : batch;
    for (;;) {
        if ((rc = launch_mixed_batch(c, batch, inner))) return rc;
        HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->h_ctl->done) break;
    }
    if (cycles) *cycles = c->h_ctl->iter;
    if (err)
        for (int q = 0; q < 5; q++) err[q] = c->h_ctl->err[q];
    int st = c->h_ctl->status;
    if (st == 3) return fail(NIM_E_MAXITER, "cycle cap reached");
    if (st == 4) return fail(NIM_E_NONFINITE, "non-finite residual");
    return NIM_OK;
}

// ---- for the acceleration driver (nim_accel.hip) ---------------------------
int32_t nim_internal_reset(nim_ctx *c, const double tol[5], int max_iter, int min_iter, bool mixed)
{
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    return reset_ctl(c, tol, max_iter, min_iter, mixed);
}
// n FP64 passes (one captured graph)
int32_t nim_internal_passes(nim_ctx *c, int n) { return launch_batch(c, n); }
// n mixed cycles of `inner` FP32 passes each (one captured graph)
int32_t nim_internal_cycles(nim_ctx *c, int n, int inner) { return launch_mixed_batch(c, n, inner); }
int32_t nim_internal_mixed_prepare(nim_ctx *c)
{
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    return mixed_prepare(c);
}
// the iteration control block after the work queued so far (synchronises
// the stream; *out points at the context's pinned copy)
int32_t nim_internal_read_ctl(nim_ctx *c, const Ctl **out)
{
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = c->h_ctl;
    return NIM_OK;
}
void **nim_internal_accel_slot(nim_ctx *c) { return &c->accel; }

extern "C" int32_t nim_residual(nim_ctx *c, double err[5])
{
    if (!c || !err) return fail(NIM_E_ARG, "null argument");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    int32_t rc = mixed_prepare(c);
    if (rc) return rc;
    const double big[5] = {1e300, 1e300, 1e300, 1e300, 1e300};
    if ((rc = reset_ctl(c, big, 1, 0))) return rc;
    launch_residual(c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 5; q++) err[q] = c->h_ctl->err[q];
    if (c->h_ctl->status == 4) return fail(NIM_E_NONFINITE, "non-finite residual");
    return NIM_OK;
}

extern "C" int32_t nim_end_step(nim_ctx *c, double *dudt, double *dTdt, double *dil)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    hipLaunchKernelGGL(k_end_step_partial, dim3(c->nred), dim3(BS), 0, c->stream, c->dev);
    hipLaunchKernelGGL(k_end_step_final, dim3(1), dim3(BS), 0, c->stream, c->dev, c->nred, c->d_out);
    HIPCHK(hipGetLastError());
    int32_t rc = d2h(c, c->h_out, c->d_out, 3 * sizeof(double));
    if (rc) return rc;
    if (dudt) *dudt = c->h_out[0];
    if (dTdt) *dTdt = c->h_out[1];
    if (dil) *dil = c->h_out[2];
    return NIM_OK;
}

extern "C" int32_t nim_end_step_noalloc(nim_ctx *c, double *out)
{
    if (!c || !out) return fail(NIM_E_ARG, "null argument");
    if (c->N > NOALLOC_MAX_CELLS)
        return fail(NIM_E_ARG, "nim_end_step_noalloc: at most 16,777,216 cells (fixed scratch size)");
    if (!c->scratch) return fail(NIM_E_HIP, "nim_end_step_noalloc: scratch address unavailable on this device");
    // `out` must be device-accessible: a kernel writes it
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, out) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NIM_E_ARG, "nim_end_step_noalloc: out is not device memory");
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != c->device)
        return fail(NIM_E_ARG, "nim_end_step_noalloc: out is not memory of the context's device");
    // the same kernels as nim_end_step, with the partials in `scratch`;
    // every block writes its partials before the final kernel reads them,
    // so nothing needs clearing
    Dev d = c->dev;
    d.red = c->scratch;
    hipLaunchKernelGGL(k_end_step_partial, dim3(c->nred), dim3(BS), 0, c->stream, d);
    hipLaunchKernelGGL(k_end_step_final, dim3(1), dim3(BS), 0, c->stream, d, c->nred, out);
    HIPCHK(hipGetLastError());
    return NIM_OK;
}

extern "C" int32_t nim_advance_time(nim_ctx *c)
{
    if (!c) return fail(NIM_E_ARG, "null context");
    hipLaunchKernelGGL(k_advance, dim3(nblk(c->N)), dim3(BS), 0, c->stream, c->dev);
    HIPCHK(hipGetLastError());
    return NIM_OK;
}

extern "C" int32_t nim_nusselt(nim_ctx *c, double out[5])
{
    if (!c || !out) return fail(NIM_E_ARG, "null argument");
    hipLaunchKernelGGL(k_nusselt, dim3(1), dim3(BS), 0, c->stream, c->dev, c->prm.R_i, c->prm.R_o, c->d_out);
    HIPCHK(hipGetLastError());
    int32_t rc = d2h(c, c->h_out, c->d_out, 5 * sizeof(double));
    if (rc) return rc;
    for (int q = 0; q < 5; q++) out[q] = c->h_out[q];
    return NIM_OK;
}

// HIP events after every launch of a timing pass, each tagged with the
// kernel class of the launch it closes
struct EventLog {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    size_t used = 0;
    bool ok = true;
    explicit EventLog(hipStream_t st) : s(st) {}
    ~EventLog()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    void start()
    {
        used = 0;
        cls.clear();
        rec(-1);
    }
    void rec(int k)
    {
        if (used == ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) { ok = false; return; }
            ev.push_back(e);
        }
        if (hipEventRecord(ev[used], s) != hipSuccess) ok = false;
        used++;
        cls.push_back(k);
    }
    // adds the time of every closed launch to ms[class]; counts launches once
    int32_t collect(double *ms, int32_t *launches, bool count)
    {
        if (!ok || used < 2) return fail(NIM_E_HIP, "event recording failed");
        HIPCHK(hipEventSynchronize(ev[used - 1]));
        for (size_t q = 1; q < used; q++) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, ev[q - 1], ev[q]));
            ms[cls[q]] += t;
            if (launches && count) launches[cls[q]]++;
        }
        return NIM_OK;
    }
};

extern "C" int32_t nim_kernel_times(nim_ctx *c, int32_t iters, double *ms, int32_t *launches)
{
    if (!c || !ms || iters < 1) return fail(NIM_E_ARG, "bad argument");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    const double tol[5] = {0, 0, 0, 0, 0};
    int32_t rc = reset_ctl(c, tol, 1 << 30, 1 << 30);
    if (rc) return rc;
    for (int q = 0; q < NIM_KERNEL_CLASSES; q++) {
        ms[q] = 0.;
        if (launches) launches[q] = 0;
    }
    EventLog log(c->stream);
    auto mark = [&](int k) { log.rec(k); };
    for (int it = 0; it < iters; it++) {
        log.start();
        for (int st = 0; st <= 6; st++) launch_stage(c, st, mark, c->fuse_T);
        hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, c->stream, c->dev);
        mark(8);
        HIPCHK(hipGetLastError());
        if ((rc = log.collect(ms, launches, it == 0))) return rc;
    }
    for (int q = 0; q < NIM_KERNEL_CLASSES; q++) ms[q] /= iters;
    // a non-finite increment stops the kernels early, which would time
    // nothing: report it
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_ctl->status == 4) return fail(NIM_E_NONFINITE, "non-finite increment while timing");
    return NIM_OK;
}

extern "C" int32_t nim_kernel_times_mixed(nim_ctx *c, int32_t cycles, int32_t inner, double *ms, int32_t *launches)
{
    if (!c || !ms || cycles < 1 || inner < 1 || inner > 1024) return fail(NIM_E_ARG, "bad argument");
    if (!c->has_geo || !c->has_pcoef) return fail(NIM_E_STATE, "geometry or pressure coefficients missing");
    int32_t rc = mixed_prepare(c);
    if (rc) return rc;
    const double tol[5] = {0, 0, 0, 0, 0};
    if ((rc = reset_ctl(c, tol, 1 << 30, 1 << 30, true))) return rc;
    for (int q = 0; q < NIM_MIXED_CLASSES; q++) {
        ms[q] = 0.;
        if (launches) launches[q] = 0;
    }
    EventLog log(c->stream);
    auto mark = [&](int k) { log.rec(k); };
    for (int it = 0; it < cycles; it++) {
        log.start();
        launch_residual(c, mark);
        for (int p = 0; p < inner; p++) launch_delta(c, mark);
        launch_update(c, inner);
        mark(11);
        HIPCHK(hipGetLastError());
        if ((rc = log.collect(ms, launches, it == 0))) return rc;
    }
    for (int q = 0; q < NIM_MIXED_CLASSES; q++) ms[q] /= cycles;
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->dev.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_ctl->status == 4) return fail(NIM_E_NONFINITE, "non-finite residual while timing");
    return NIM_OK;
}

// ---- ensembles -----------------------------------------------------------
// launch shapes: blocks per case of each kind of launch
enum { SH_CELL64, SH_CELL256, SH_COL64, SH_COL256, SH_WALL, SH_UPD, SH_NRED, NSHAPE };

struct nim_ens {
    int n = 0;
    std::vector<nim_ctx *> cs;
    hipStream_t stream = nullptr;
    Dev *d_dev = nullptr;
    View<double> *d_v64 = nullptr;
    View<float> *d_v32 = nullptr;
    int *d_bstart = nullptr;         // [NSHAPE][n + 1]
    int total[NSHAPE] = {};
    Ctl *d_ctl = nullptr, *h_ctl = nullptr;
    bool views = false;              // mixed views uploaded
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gbatch = 0, ginner = -1;
    // whole-step form (nim_ens_time_step): per-step kernels of all cases in
    // two captured graphs around the solve, results of all cases in one read
    int *d_nred = nullptr;           // [n] partial blocks of each case's end-of-step reduction
    double *d_R = nullptr;           // [2n] R_i, R_o
    double *d_out = nullptr, *h_out = nullptr;  // [8n] dudt, dTdt, dilatation, Nu[5]
    hipGraph_t g_pre[2] = {nullptr, nullptr}, g_post = nullptr;
    hipGraphExec_t x_pre[2] = {nullptr, nullptr}, x_post = nullptr;
    std::vector<void *> allocs;
};

static Ens ens_args(const nim_ens *e, int shape)
{
    Ens a;
    a.dev = e->d_dev;
    a.v64 = e->d_v64;
    a.v32 = e->d_v32;
    a.bstart = e->d_bstart + (size_t)shape * (e->n + 1);
    a.ncase = e->n;
    return a;
}

extern "C" int32_t nim_ens_destroy(nim_ens *e)
{
    if (!e) return NIM_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (int q = 0; q < 2; q++) {
        if (e->x_pre[q]) (void)hipGraphExecDestroy(e->x_pre[q]);
        if (e->g_pre[q]) (void)hipGraphDestroy(e->g_pre[q]);
    }
    if (e->x_post) (void)hipGraphExecDestroy(e->x_post);
    if (e->g_post) (void)hipGraphDestroy(e->g_post);
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->h_ctl) (void)hipHostFree(e->h_ctl);
    if (e->h_out) (void)hipHostFree(e->h_out);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return NIM_OK;
}

extern "C" int32_t nim_ens_create(nim_ctx *const *cases, int32_t ncase, nim_ens **out)
{
    if (!cases || !out || ncase < 1 || ncase > 4096) return fail(NIM_E_ARG, "bad ensemble arguments");
    *out = nullptr;
    for (int q = 0; q < ncase; q++) {
        if (!cases[q]) return fail(NIM_E_ARG, "null case context");
        if (!cases[q]->has_geo || !cases[q]->has_pcoef)
            return fail(NIM_E_STATE, "every case needs geometry and pressure coefficients");
    }
    nim_ens *e = new nim_ens();
    e->n = ncase;
    e->cs.assign(cases, cases + ncase);
    auto bad = [&](int32_t rc) { nim_ens_destroy(e); return rc; };
    std::vector<Dev> devs(ncase);
    std::vector<int> bs((size_t)NSHAPE * (ncase + 1));
    for (int sh = 0; sh < NSHAPE; sh++) {
        int acc = 0;
        for (int q = 0; q < ncase; q++) {
            const nim_ctx *c = cases[q];
            bs[(size_t)sh * (ncase + 1) + q] = acc;
            int cc = (c->NT / 2) * ((c->NR + 1) / 2) * c->NZ;
            switch (sh) {
            case SH_CELL64: acc += cell_grid(c, 64); break;
            case SH_CELL256: acc += cell_grid(c, BS); break;
            case SH_COL64: acc += (cc + 63) / 64; break;
            case SH_COL256: acc += nblk(cc); break;
            case SH_WALL: acc += nblk(2 * c->NT * c->NR); break;
            case SH_NRED: acc += c->nred; break;  // blocks of 256 cells in storage order
            default: acc += nblk(std::max(c->N, c->NF)); break;
            }
        }
        bs[(size_t)sh * (ncase + 1) + ncase] = acc;
        e->total[sh] = acc;
    }
    for (int q = 0; q < ncase; q++) devs[q] = cases[q]->dev;
    auto alloc = [&](void **p, size_t bytes) -> int32_t {
        HIPCHK(hipMalloc(p, bytes));
        e->allocs.push_back(*p);
        return NIM_OK;
    };
    int32_t rc;
    if ((rc = alloc((void **)&e->d_dev, sizeof(Dev) * ncase)) ||
        (rc = alloc((void **)&e->d_v64, sizeof(View<double>) * ncase)) ||
        (rc = alloc((void **)&e->d_v32, sizeof(View<float>) * ncase)) ||
        (rc = alloc((void **)&e->d_bstart, sizeof(int) * bs.size())) ||
        (rc = alloc((void **)&e->d_ctl, sizeof(Ctl) * ncase)) ||
        (rc = alloc((void **)&e->d_nred, sizeof(int) * ncase)) ||
        (rc = alloc((void **)&e->d_R, sizeof(double) * 2 * ncase)) ||
        (rc = alloc((void **)&e->d_out, sizeof(double) * 8 * ncase)))
        return bad(rc);
    if (hipHostMalloc((void **)&e->h_ctl, sizeof(Ctl) * ncase, 0) != hipSuccess ||
        hipHostMalloc((void **)&e->h_out, sizeof(double) * 8 * ncase, 0) != hipSuccess)
        return bad(fail(NIM_E_HIP, "hipHostMalloc failed"));
    {
        std::vector<int> nred(ncase);
        std::vector<double> R(2 * (size_t)ncase);
        for (int q = 0; q < ncase; q++) {
            nred[q] = cases[q]->nred;
            R[2 * q] = cases[q]->prm.R_i;
            R[2 * q + 1] = cases[q]->prm.R_o;
        }
        if (hipMemcpy(e->d_nred, nred.data(), sizeof(int) * ncase, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(e->d_R, R.data(), sizeof(double) * 2 * ncase, hipMemcpyHostToDevice) != hipSuccess)
            return bad(fail(NIM_E_HIP, "ensemble table upload failed"));
    }
    if (hipStreamCreate(&e->stream) != hipSuccess) return bad(fail(NIM_E_HIP, "hipStreamCreate failed"));
    if (hipMemcpy(e->d_dev, devs.data(), sizeof(Dev) * ncase, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_bstart, bs.data(), sizeof(int) * bs.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(NIM_E_HIP, "ensemble table upload failed"));
    *out = e;
    return NIM_OK;
}

// one FP64 pass of every case
static void ens_pass(nim_ens *e)
{
    hipStream_t s = e->stream;
    hipLaunchKernelGGL(k_e_mom_src, dim3(e->total[SH_CELL256]), dim3(BS), 0, s, ens_args(e, SH_CELL256));
    for (int q = 0; q < 4; q++)
        hipLaunchKernelGGL(k_e_mom_sweep, dim3(e->total[SH_COL256]), dim3(BS), 0, s, ens_args(e, SH_COL256), q);
    hipLaunchKernelGGL(k_e_ppe_src, dim3(e->total[SH_CELL256]), dim3(BS), 0, s, ens_args(e, SH_CELL256));
    for (int q = 0; q < 4; q++)
        hipLaunchKernelGGL(k_e_ppe_sweep, dim3(e->total[SH_COL256]), dim3(BS), 0, s, ens_args(e, SH_COL256), q);
    hipLaunchKernelGGL(k_e_T_coef, dim3(e->total[SH_CELL256]), dim3(BS), 0, s, ens_args(e, SH_CELL256));
    hipLaunchKernelGGL(k_e_T_src, dim3(e->total[SH_CELL256]), dim3(BS), 0, s, ens_args(e, SH_CELL256));
    for (int q = 0; q < 4; q++)
        hipLaunchKernelGGL(k_e_T_sweep, dim3(e->total[SH_COL256]), dim3(BS), 0, s, ens_args(e, SH_COL256), q);
    hipLaunchKernelGGL(k_e_check, dim3(e->n), dim3(64), 0, s, ens_args(e, SH_CELL256));
}

// one mixed cycle of every case: FP64 residual, `inner` FP32 passes, update
static void ens_cycle(nim_ens *e, int inner)
{
    hipStream_t s = e->stream;
    const Ens c64 = ens_args(e, SH_CELL64), c256 = ens_args(e, SH_CELL256);
    const Ens k64 = ens_args(e, SH_COL64), k256 = ens_args(e, SH_COL256);
    const dim3 g64(e->total[SH_CELL64]), g256(e->total[SH_CELL256]), gk64(e->total[SH_COL64]), gk256(e->total[SH_COL256]);
    hipLaunchKernelGGL(k_e_mom_src, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_res_mom, g64, BS3, 0, s, c64);
    hipLaunchKernelGGL(k_e_res_ppe_src, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_res_ppe, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_res_T_coef, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_T_src, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_res_T, g256, dim3(BS), 0, s, c256);
    hipLaunchKernelGGL(k_e_check, dim3(e->n), dim3(64), 0, s, c256);
    for (int p = 0; p < inner; p++) {
        hipLaunchKernelGGL(k_e_d_mom_src, g256, dim3(BS), 0, s, c256);
        for (int q = 0; q < 4; q++) hipLaunchKernelGGL(k_e_d_mom_sweep, gk64, BS3, 0, s, k64, q);
        hipLaunchKernelGGL(k_e_d_ppe_src, g256, dim3(BS), 0, s, c256);
        for (int q = 0; q < 4; q++) hipLaunchKernelGGL(k_e_d_ppe_sweep, gk256, dim3(BS), 0, s, k256, q);
        hipLaunchKernelGGL(k_e_d_zero_flux, dim3(e->total[SH_WALL]), dim3(BS), 0, s, ens_args(e, SH_WALL));
        hipLaunchKernelGGL(k_e_d_T_src, g256, dim3(BS), 0, s, c256);
        for (int q = 0; q < 4; q++) hipLaunchKernelGGL(k_e_d_T_sweep, gk256, dim3(BS), 0, s, k256, q);
    }
    hipLaunchKernelGGL(k_e_mix_update, dim3(e->total[SH_UPD]), dim3(BS), 0, s, ens_args(e, SH_UPD), 1. / inner);
}

static int32_t ens_batch(nim_ens *e, int batch, int inner)
{
    auto body = [&]() {
        for (int b = 0; b < batch; b++) {
            if (inner > 0) ens_cycle(e, inner);
            else ens_pass(e);
        }
    };
    if (!use_graph()) {
        body();
        HIPCHK(hipGetLastError());
        return NIM_OK;
    }
    if (!e->gexec || e->gbatch != batch || e->ginner != inner) {
        if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
        if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        body();
        HIPCHK(hipStreamEndCapture(e->stream, &e->graph));
        HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        e->gbatch = batch;
        e->ginner = inner;
    }
    HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    return NIM_OK;
}

static int32_t ens_upload_views(nim_ens *e)
{
    if (e->views) return NIM_OK;
    std::vector<View<double>> v64(e->n);
    std::vector<View<float>> v32(e->n);
    for (int q = 0; q < e->n; q++) {
        v64[q] = e->cs[q]->mx.v64;
        v32[q] = e->cs[q]->mx.v32;
    }
    HIPCHK(hipMemcpy(e->d_v64, v64.data(), sizeof(View<double>) * e->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_v32, v32.data(), sizeof(View<float>) * e->n, hipMemcpyHostToDevice));
    e->views = true;
    return NIM_OK;
}

// the passes (inner = 0) or mixed cycles of every case on the ensemble
// stream until all have stopped; control blocks already reset
static int32_t ens_loop(nim_ens *e, int inner, int32_t *iters, double *err)
{
    int32_t rc;
    int batch = inner > 0 ? std::max(1, (16 + inner) / (inner + 1)) : batch_size();
    for (;;) {
        if ((rc = ens_batch(e, batch, inner))) return rc;
        hipLaunchKernelGGL(k_e_gather, dim3(1), dim3(64), 0, e->stream, ens_args(e, SH_CELL256), e->d_ctl);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(e->h_ctl, e->d_ctl, sizeof(Ctl) * e->n, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        bool all = true;
        for (int q = 0; q < e->n; q++) all = all && e->h_ctl[q].done;
        if (all) break;
    }
    int32_t worst = NIM_OK;
    for (int q = 0; q < e->n; q++) {
        const Ctl &t = e->h_ctl[q];
        // keep each case's host copy of its control block current
        *e->cs[q]->h_ctl = t;
        if (iters) iters[q] = t.iter;
        if (err)
            for (int f = 0; f < 5; f++) err[(size_t)q * 5 + f] = t.err[f];
        if (t.status == 4) worst = NIM_E_NONFINITE;
        else if (t.status == 3 && worst == NIM_OK) worst = NIM_E_MAXITER;
    }
    if (worst == NIM_E_NONFINITE) return fail(worst, "non-finite increment in a case");
    if (worst == NIM_E_MAXITER) return fail(worst, "iteration cap reached in a case");
    return NIM_OK;
}

extern "C" int32_t nim_ens_inner_solve(nim_ens *e, const double tol[5], int32_t max_iter, int32_t min_iter,
                                       int32_t inner, int32_t *iters, double *err)
{
    if (!e || !tol) return fail(NIM_E_ARG, "null argument");
    if (max_iter < 1 || inner < 0 || inner > 1024) return fail(NIM_E_ARG, "bad iteration arguments");
    int32_t rc;
    for (nim_ctx *c : e->cs) {
        if (inner > 0 && (rc = mixed_prepare(c))) return rc;
        if ((rc = reset_ctl(c, tol, max_iter, min_iter))) return rc;
    }
    if (inner > 0 && (rc = ens_upload_views(e))) return rc;
    // the cases' own streams carry their per-step work (coefficients,
    // casts, control reset): finish it before the ensemble stream starts
    for (nim_ctx *c : e->cs) HIPCHK(hipStreamSynchronize(c->stream));
    return ens_loop(e, inner, iters, err);
}

// capture `body` (launches on the ensemble stream) into g / x once
template <class Body>
static int32_t ens_graph(nim_ens *e, hipGraph_t &g, hipGraphExec_t &x, Body body)
{
    if (x) return NIM_OK;
    if (!use_graph()) {
        body();
        HIPCHK(hipGetLastError());
        return NIM_OK;
    }
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    body();
    HIPCHK(hipStreamEndCapture(e->stream, &g));
    HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    return NIM_OK;
}

extern "C" int32_t nim_ens_time_step(nim_ens *e, const double tol[5], int32_t max_iter, int32_t min_iter,
                                     int32_t inner, int32_t *iters, double *err, double *out)
{
    if (!e || !tol) return fail(NIM_E_ARG, "null argument");
    if (max_iter < 1 || inner < 0 || inner > 1024) return fail(NIM_E_ARG, "bad iteration arguments");
    int32_t rc;
    // the cases' streams must be idle; the mixed loop's buffers and FP32
    // constants exist (set up once, per case)
    for (nim_ctx *c : e->cs) {
        if (inner > 0 && (!c->mx.alloc || c->mx.version != c->const_version) && (rc = mixed_prepare(c))) return rc;
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (inner > 0 && (rc = ens_upload_views(e))) return rc;
    hipStream_t s = e->stream;
    // control blocks, then the momentum coefficients of every case (with
    // their FP32 copies for the mixed loop): one captured graph
    for (nim_ctx *c : e->cs)
        hipLaunchKernelGGL(k_solve_reset, dim3(1), dim3(256), 0, s, c->dev, tol[0], tol[1], tol[2], tol[3], tol[4],
                           max_iter, min_iter, c->p_measure);
    HIPCHK(hipGetLastError());
    const int cast = inner > 0;
    auto pre = [&]() {
        hipLaunchKernelGGL(k_e_coef_A, dim3(e->total[SH_NRED]), dim3(BS), 0, s, ens_args(e, SH_NRED));
        hipLaunchKernelGGL(k_e_coef_F, dim3(e->total[SH_NRED]), dim3(BS), 0, s, ens_args(e, SH_NRED), cast);
    };
    if ((rc = ens_graph(e, e->g_pre[cast], e->x_pre[cast], pre))) return rc;
    if (use_graph()) HIPCHK(hipGraphLaunch(e->x_pre[cast], s));
    if ((rc = ens_loop(e, inner, iters, err))) return rc;
    // end of step, time shift and Nusselt numbers of every case: one graph
    auto post = [&]() {
        hipLaunchKernelGGL(k_e_end_step_partial, dim3(e->total[SH_NRED]), dim3(BS), 0, s, ens_args(e, SH_NRED));
        hipLaunchKernelGGL(k_e_end_step_final, dim3(e->n), dim3(BS), 0, s, ens_args(e, SH_NRED),
                           (const int *)e->d_nred, e->d_out);
        hipLaunchKernelGGL(k_e_advance, dim3(e->total[SH_NRED]), dim3(BS), 0, s, ens_args(e, SH_NRED));
        hipLaunchKernelGGL(k_e_nusselt, dim3(e->n), dim3(BS), 0, s, ens_args(e, SH_NRED), (const double *)e->d_R,
                           e->d_out);
    };
    if ((rc = ens_graph(e, e->g_post, e->x_post, post))) return rc;
    if (use_graph()) HIPCHK(hipGraphLaunch(e->x_post, s));
    HIPCHK(hipMemcpyAsync(e->h_out, e->d_out, sizeof(double) * 8 * e->n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (out) std::memcpy(out, e->h_out, sizeof(double) * 8 * e->n);
    return NIM_OK;
}
