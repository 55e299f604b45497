// Checkpoints of the time-step state on the device and the binomial
// (revolve) plan that reverses a run of time steps with a fixed number of
// them. No counterpart in the reference (its restart files, restart.f90,
// hold one state on disk); this is the store an adjoint of the time loop
// uses to visit the states of a forward run in reverse order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "nim_hip.h"
#include "nim_kernels.h"

using namespace nim;

// shared with nim_api.hip
int32_t nim_internal_fail(int32_t code, const char *msg);
void *nim_internal_stream(nim_ctx *c);
const Dev *nim_internal_dev(nim_ctx *c);

// ---- plan ----------------------------------------------------------------
// cost(l, s): fewest forward steps that reverse l steps when the state at
// the start of the range is current and stored, and s more snapshots are
// free. One step is reversed in place; with no free snapshot every step k
// is reached again from the start (l (l - 1) / 2 steps); otherwise advance
// m steps, store, reverse the right part with one snapshot fewer, restore
// the start and reverse the left part. With c snapshots in all (the start
// included) and l = C(c + r, c) this is Griewank's optimum
// r l - C(c + r, c + 1).
namespace {
struct Planner {
    int N, S;
    std::vector<long long> cost;  // [(l) * (S + 1) + s]
    std::vector<int> split;
    std::vector<int32_t> code, arg;
    std::vector<int> free_slots;
    long long fwd = 0;

    long long &C(int l, int s) { return cost[(size_t)l * (S + 1) + s]; }
    int &M(int l, int s) { return split[(size_t)l * (S + 1) + s]; }

    void tables()
    {
        cost.assign((size_t)(N + 1) * (S + 1), 0);
        split.assign((size_t)(N + 1) * (S + 1), 0);
        for (int l = 1; l <= N; l++)
            for (int s = 0; s <= S; s++) {
                if (l == 1) { C(l, s) = 0; continue; }
                if (s == 0) { C(l, s) = (long long)l * (l - 1) / 2; continue; }
                long long best = -1;
                int bm = 1;
                for (int m = 1; m < l; m++) {
                    long long v = m + C(l - m, s - 1) + C(m, s);
                    if (best < 0 || v < best) { best = v; bm = m; }
                }
                C(l, s) = best;
                M(l, s) = bm;
            }
    }
    void emit(int32_t c, int32_t a)
    {
        code.push_back(c);
        arg.push_back(a);
        if (c == NIM_RV_ADVANCE) fwd += a;
    }
    // state a is current and held in slot `home`; s snapshots free
    void plan(int a, int b, int s, int home)
    {
        int l = b - a;
        if (l == 1) { emit(NIM_RV_REVERSE, a); return; }
        if (s == 0) {
            for (int k = b - 1; k >= a; k--) {
                if (k > a) emit(NIM_RV_ADVANCE, k - a);
                emit(NIM_RV_REVERSE, k);
                if (k > a) emit(NIM_RV_RESTORE, home);
            }
            return;
        }
        int m = M(l, s);
        int slot = free_slots.back();
        free_slots.pop_back();
        emit(NIM_RV_ADVANCE, m);
        emit(NIM_RV_STORE, slot);
        plan(a + m, b, s - 1, slot);
        free_slots.push_back(slot);
        emit(NIM_RV_RESTORE, home);
        plan(a, a + m, s, home);
    }
};
}  // namespace

extern "C" int32_t nim_revolve_plan(int32_t nsteps, int32_t nslots, int32_t *codes, int32_t *args, int64_t capacity,
                                    int64_t *n_actions, int64_t *forward_steps)
{
    if (nsteps < 1 || nsteps > 4096 || nslots < 1 || nslots > 256)
        return nim_internal_fail(NIM_E_ARG, "nsteps must be in [1, 4096] and nslots in [1, 256]");
    Planner p;
    p.N = nsteps;
    p.S = nslots - 1;
    p.tables();
    for (int q = nslots - 1; q >= 1; q--) p.free_slots.push_back(q);
    p.emit(NIM_RV_STORE, 0);
    p.plan(0, nsteps, nslots - 1, 0);
    if (n_actions) *n_actions = (int64_t)p.code.size();
    if (forward_steps) *forward_steps = p.fwd;
    if (codes && args) {
        if ((int64_t)p.code.size() > capacity) return nim_internal_fail(NIM_E_ARG, "action buffer too small");
        std::copy(p.code.begin(), p.code.end(), codes);
        std::copy(p.arg.begin(), p.arg.end(), args);
    }
    return NIM_OK;
}

// ---- device snapshots -----------------------------------------------------
struct nim_ckpt {
    nim_ctx *ctx = nullptr;
    int nslots = 0;
    size_t cells = 0, faces = 0;  // doubles per snapshot of each block
    double *pool = nullptr;       // [nslots][cells + faces]
};

extern "C" int32_t nim_ckpt_create(nim_ctx *ctx, int32_t nslots, nim_ckpt **out)
{
    if (!ctx || !out || nslots < 1 || nslots > 4096) return nim_internal_fail(NIM_E_ARG, "bad checkpoint arguments");
    *out = nullptr;
    const Dev *d = nim_internal_dev(ctx);
    nim_ckpt *k = new nim_ckpt();
    k->ctx = ctx;
    k->nslots = nslots;
    k->cells = (size_t)NSC * d->N;
    k->faces = (size_t)5 * d->NF;
    if (hipMalloc((void **)&k->pool, sizeof(double) * (k->cells + k->faces) * nslots) != hipSuccess) {
        delete k;
        return nim_internal_fail(NIM_E_HIP, "checkpoint pool allocation failed");
    }
    *out = k;
    return NIM_OK;
}

extern "C" int32_t nim_ckpt_destroy(nim_ckpt *k)
{
    if (!k) return NIM_OK;
    (void)hipStreamSynchronize((hipStream_t)nim_internal_stream(k->ctx));
    (void)hipFree(k->pool);
    delete k;
    return NIM_OK;
}

static int32_t ckpt_copy(nim_ckpt *k, int32_t slot, bool store)
{
    if (!k || slot < 0 || slot >= k->nslots) return nim_internal_fail(NIM_E_ARG, "checkpoint slot out of range");
    const Dev *d = nim_internal_dev(k->ctx);
    hipStream_t s = (hipStream_t)nim_internal_stream(k->ctx);
    double *base = k->pool + (size_t)slot * (k->cells + k->faces);
    hipError_t e1, e2;
    if (store) {
        e1 = hipMemcpyAsync(base, d->sc, sizeof(double) * k->cells, hipMemcpyDeviceToDevice, s);
        e2 = hipMemcpyAsync(base + k->cells, d->fc, sizeof(double) * k->faces, hipMemcpyDeviceToDevice, s);
    } else {
        e1 = hipMemcpyAsync(d->sc, base, sizeof(double) * k->cells, hipMemcpyDeviceToDevice, s);
        e2 = hipMemcpyAsync(d->fc, base + k->cells, sizeof(double) * k->faces, hipMemcpyDeviceToDevice, s);
    }
    if (e1 != hipSuccess || e2 != hipSuccess) return nim_internal_fail(NIM_E_HIP, "checkpoint copy failed");
    return NIM_OK;
}

extern "C" int32_t nim_ckpt_store(nim_ckpt *k, int32_t slot) { return ckpt_copy(k, slot, true); }
extern "C" int32_t nim_ckpt_restore(nim_ckpt *k, int32_t slot) { return ckpt_copy(k, slot, false); }
